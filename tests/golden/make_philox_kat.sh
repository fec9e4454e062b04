#!/bin/sh
# Regenerates philox_kat.json from rocRAND's Philox4x32-10 (host code only).
set -e
here=$(cd "$(dirname "$0")" && pwd)
hipcc -O1 -o /tmp/gen_philox_kat "$here/gen_philox_kat.cpp"
/tmp/gen_philox_kat > "$here/philox_kat.json"
