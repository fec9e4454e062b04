"""Writes the ANO fixture from the reference's own data files (data, not code).

Input: data process/ANO_Tnum/data_record_T_ANO_100.txt (18 runs x 100 tasks x
(21 bests + initial worst), the record format of proposed model/SG-EMTF_DE.cu:1793-1811)
and data_record_T_ANO_100_1677482963.txt, the ANO curve the reference's program
(data process/源.cpp:122-205) printed for it.  Output: ano_T100_records.npz
(float32 [run, task, point]) and ano_T100_expected.txt (the printed curve, verbatim).

Run once in a checkout that has the reference under /root/reference.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/data process/ANO_Tnum/"


def main():
    sys.path.insert(0, os.path.join(HERE, "..", "..", "sg-emtf_amd"))
    from emtf.convergence import read_record
    R = read_record(SRC + "data_record_T_ANO_100.txt")
    np.savez_compressed(os.path.join(HERE, "ano_T100_records.npz"), records=R)
    with open(SRC + "data_record_T_ANO_100_1677482963.txt") as fh:
        txt = fh.read()
    with open(os.path.join(HERE, "ano_T100_expected.txt"), "w") as fh:
        fh.write(txt)
    print(R.shape, len(txt))


if __name__ == "__main__":
    main()
