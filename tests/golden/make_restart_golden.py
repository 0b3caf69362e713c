"""Write tests/golden/restart_dyn_small.bin with the reference FIM restart
writer (write_restart_dyn, restart_dyn.F90, built into oracle/_ref/bin/restref_r4
by oracle/build_ref.sh) from the seeded contents of restart_values.make.
Run from the repository root: python tests/golden/make_restart_golden.py"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, HERE)

from oracle import refrun  # noqa: E402
import restart_values  # noqa: E402

NVL, NIP, NTRA, NTRB = 4, 12, 2, 1

if __name__ == "__main__":
    rd = restart_values.make(NVL, NIP, NTRA, NTRB)
    data = refrun.run_restref_write(restart_values.value_stream(rd, NTRB))
    with open(os.path.join(HERE, "restart_dyn_small.bin"), "wb") as f:
        f.write(data)
    print("wrote restart_dyn_small.bin, %d bytes" % len(data))
