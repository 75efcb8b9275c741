"""Copies the h10 integral / trial file and the first block of its reference
estimates out of ipie's test data (ipie/qmc/tests/reference_data/
h10_cc-pvtz_batched/{afqmc.h5,reference.json}).  Block 0 is the estimate of
walkers at the trial (deterministic); later blocks depend on ipie's random
stream over 4 MPI ranks and are not kept.

    python tests/golden/make_h10_fixture.py
"""

import json
import os
import shutil

SRC = "/root/reference/ipie/qmc/tests/reference_data/h10_cc-pvtz_batched"
HERE = os.path.dirname(os.path.abspath(__file__))

if __name__ == "__main__":
    shutil.copyfile(os.path.join(SRC, "afqmc.h5"), os.path.join(HERE, "h10_afqmc.h5"))
    with open(os.path.join(SRC, "reference.json")) as fh:
        ref = json.load(fh)
    keep = {k: ref[k][0] for k in ("ETotal", "E1Body", "E2Body", "Weight", "Overlap")}
    keep["nwalkers_total"] = ref["Weight"][0]
    with open(os.path.join(HERE, "h10_block0.json"), "w") as fh:
        json.dump(keep, fh, indent=1)
    print(keep)
