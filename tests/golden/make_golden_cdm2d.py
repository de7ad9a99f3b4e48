#!/usr/bin/env python3
"""Writes tests/golden/cdm2d_vectors.npz from the compiled reference
(oracle/_ref/libtiplref.so, built by `make -C oracle ref` from the reference
headers): seeded inputs and the reference's outputs for every step of the
2-D CDM (cdm<out,uint8,2>) and one whole 2-D cdm() run.  test_oracle_golden.py
compares the oracle with them where the reference build is not available.

Run once here: python tests/golden/make_golden_cdm2d.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from oracle.bindings import Reference  # noqa: E402
from test_oracle_cdm2d import field2, image2  # noqa: E402


def main():
    ref = Reference()
    out = {}
    v = np.random.default_rng(201).integers(0, 256, (17, 23)).astype(np.uint8)
    out.update(down_in=v, down_out=ref.downsample2(v))
    f = np.random.default_rng(202).standard_normal((7, 9, 2)).astype(np.float32)
    out.update(up_in=f, up_out=ref.upsample2(f, (14, 18)))

    shape = (22, 26)
    Is = image2(shape, 203, background=False)
    It = image2(shape, 204)
    d = field2(shape, 205, 3.0)
    Js = ref.compose2(Is, d)
    out.update(Is=Is, It=It, d=d, Js=Js)
    gr, cr = ref.gradient2(Js, It, "r")
    gd, cd = ref.gradient2(Js, It, "d")
    out.update(grad_r=gr, cost_r=np.float32(cr), grad_d=gd, cost_d=np.float32(cd))
    nd = field2(shape, 206, 1.0, 0.0)
    ps = ref.poisson2(nd)
    out.update(nd=nd, poisson=ps, maxlen=np.float32(ref.max_length2(ps)))
    out["accumulate"] = ref.accumulate2(d, ps)
    out["smooth"] = ref.smooth2(ps, d, 0.4)

    from scipy.ndimage import shift
    It2 = image2((60, 52), 207, sigma=2.0)
    Is2 = shift(It2.astype(np.float32), (1.5, -2.0), order=1).astype(np.uint8)
    ok, f2, _ = ref.cdm2(It2, Is2)
    out.update(full_It=It2, full_Is=Is2, full_ok=np.int32(ok), full_field=f2)

    path = os.path.join(HERE, "cdm2d_vectors.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    main()
