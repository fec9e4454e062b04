"""Builds libemtf.so (HIP kernels + C ABI) in-tree for gfx950.

island.hip is compiled once per (gene-block count, algorithm family) pair of the
LDS-resident kernel and once per (row-tile count, algorithm family) pair of the
wide kernel; each object holds one kernel that dispatches per task to an
objective-specialised body.  The objects are linked with the host runtime into
sg-emtf_amd/libemtf.so.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libemtf.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# -ffp-contract=off: the arithmetic contract forbids implicit fma contraction.
# -packed-fp32-ops: keep the device compiler away from v_pk_{add,mul,fma}_f32.
COMMON = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-ffp-contract=off", "-fPIC",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall", "-Wno-unused-function"]
# (gene blocks per parity half KB: dims <= 8 KB; algorithm family: 0 = DE, 2 = GA)
VARIANTS = [(kb, al) for kb in (2, 4, 7, 8, 12, 13, 16) for al in (0, 2)]
# wide kernel: (32-row tiles NT: dims <= 32 NT, rows in global memory; algorithm family)
WIDE = [(nt, al) for nt in (2, 4, 5, 6, 7, 8) for al in (0, 2)]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = [l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l and l.strip()]
    return r.returncode, "\n".join(msgs)


def build(verbose=False, jobs=None, out=None, extra=()):
    """Compile every variant and link LIB (or `out`); `extra` adds compiler flags."""
    out = out or LIB
    os.makedirs(OBJ, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 1)
    units = [([HIPCC] + COMMON + list(extra) + [f"-DEMTF_KB={kb}", f"-DEMTF_ALGO={al}", "-c",
                                  "-o", os.path.join(OBJ, f"island_kb{kb}_a{al}.o"), os.path.join(CSRC, "island.hip")])
             for kb, al in VARIANTS]
    units += [([HIPCC] + COMMON + list(extra) + [f"-DEMTF_WIDE_NT={nt}", f"-DEMTF_ALGO={al}", "-c",
                                   "-o", os.path.join(OBJ, f"island_w{nt}_a{al}.o"), os.path.join(CSRC, "island.hip")])
              for nt, al in WIDE]
    for name in ("emtf_host.cpp", "cvrp_host.cpp", "cvrp.hip", "export.hip"):
        units.append([HIPCC] + COMMON + ["-c", "-o", os.path.join(OBJ, name.split(".")[0] + ".o"), os.path.join(CSRC, name)])
    with ThreadPoolExecutor(jobs) as ex:
        results = list(ex.map(_run, units))
    for (rc, msg), cmd in zip(results, units):
        if rc != 0:
            sys.stderr.write(msg + "\n")
            raise RuntimeError("hipcc failed: " + " ".join(cmd[-3:]))
        if verbose and msg:
            print(msg)
    objs = [os.path.join(OBJ, f"island_kb{kb}_a{al}.o") for kb, al in VARIANTS] + \
        [os.path.join(OBJ, f"island_w{nt}_a{al}.o") for nt, al in WIDE] + \
        [os.path.join(OBJ, n + ".o") for n in ("emtf_host", "cvrp_host", "cvrp", "export")]
    rc, msg = _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", out + ".tmp"] + objs)
    if rc != 0:
        sys.stderr.write(msg + "\n")
        raise RuntimeError("link of libemtf.so failed")
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
