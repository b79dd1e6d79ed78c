"""Write tests/golden/carbon_sphere.npz from the reference's own test files.

Data only: tests/files/carbon_sphere.xyz (1,985 carbon atoms) and
tests/files/carbon_sphere_waxs.dat (20,000 q points, "calculated with older
version of AUSAXS, known to be accurate"), used by the reference's
feature/other/wide_angle_scattering.cpp:14-37 with 0.1 A bins, exv None and a
1 % tolerance. Every 10th q point is kept.

    python tests/golden/make_carbon_sphere.py /root/reference
"""
import sys
from pathlib import Path

import numpy as np

ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference") / "tests" / "files"
lines = (ref / "carbon_sphere.xyz").read_text().splitlines()
n = int(lines[0].split()[0])
xyz = np.array([[float(v) for v in ln.split()[1:4]] for ln in lines[2 : 2 + n]], dtype=np.float64)
assert xyz.shape == (n, 3)
dat = np.loadtxt(ref / "carbon_sphere_waxs.dat", skiprows=2)
q, I = dat[::10, 0], dat[::10, 1]
out = Path(__file__).resolve().parent / "carbon_sphere.npz"
np.savez_compressed(out, xyz=xyz, q=q, I=I)
print(f"wrote {out}: {n} atoms, {len(q)} q points (q {q[0]:g}..{q[-1]:g})")
