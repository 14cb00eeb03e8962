"""Writes tests/golden/rca_cfg0_tiny.npz: FP64-oracle outputs of a small
configuration-0-like case (64 column elements, 11 row tilts -10..10 deg,
Lanczos a=3, 6x5x12 voxels, 384 samples, 30 seeded scatterers).  The inputs
are regenerated from the seeds by simu_rca_ab_amd.synth, so only outputs are
stored.  Run: python tests/golden/make_golden.py
"""
from pathlib import Path
import sys

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

from simu_rca_ab_amd import configs as C, synth  # noqa: E402

PATH = Path(__file__).resolve().parent / "rca_cfg0_tiny.npz"
CFG = C.scaled(C.get("cfg0"), nx=6, ny=5, nz=12, n_sample=384, n_scat=30,
               scat_box=(3e-3, 3e-3, 4e-3, 8e-3))


def inputs():
    iq = synth.frame(CFG)[0].numpy()
    h = CFG.halves[0]
    return iq, C.geometry(CFG, h), C.transmit(h), C.param(CFG), C.options(CFG)


def compute():
    from oracle import rca_oracle as O
    iq, G, T, P, Op = inputs()
    dataBf, mask = O.bf(synth.to_matlab_layout(iq), G, T, P, Op)
    return {"dataBf_re": dataBf.real.ravel(order="F"), "dataBf_im": dataBf.imag.ravel(order="F"),
            "mask": mask.ravel(order="F")}


if __name__ == "__main__":
    np.savez_compressed(PATH, **compute())
    print("wrote", PATH)
