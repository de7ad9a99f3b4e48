#!/usr/bin/env python3
"""Writes tests/golden/reference_vectors.npz from the compiled reference
(oracle/_ref/libtiplref.so, built by `make -C oracle ref` from the
reference headers).  The file holds seeded inputs and the reference's
outputs for every function on the registration path; tests compare the
oracle (and through it the GPU path) against them where the reference
build is not available.

Run once here: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from conftest import prep_images, random_params, smooth_volume  # noqa: E402
from oracle.bindings import Reference  # noqa: E402


def field(shape, seed, scale=2.0, zero_frac=0.2):
    from scipy.ndimage import gaussian_filter
    r = np.random.default_rng(seed)
    d = gaussian_filter(r.standard_normal(shape + (3,)), (2, 2, 2, 0)).astype(np.float32) * scale
    d[r.random(shape) < zero_frac] = 0.0
    return d


def main():
    ref = Reference()
    out = {}
    fshape, tshape = (12, 14, 16), (13, 15, 11)
    A = smooth_volume(fshape, 101)
    B = smooth_volume(tshape, 102)
    P = random_params(np.random.default_rng(103), 6)
    fvs = np.array([1.0, 1.1, 0.9], np.float32)
    tvs = np.array([1.2, 1.0, 1.0], np.float32)
    Ts = np.stack([ref.affine_matrix(p, np.array(fshape[::-1], np.int32), fvs,
                                     np.array(tshape[::-1], np.int32), tvs) for p in P])
    out.update(aff_A=A, aff_B=B, aff_params=P, aff_fvs=fvs, aff_tvs=tvs, aff_T=Ts)
    out["aff_Tinv"] = np.stack([ref.matrix_inverse(T)[0] for T in Ts])
    out["aff_resample"] = np.stack([ref.resample(B, fshape, T) for T in Ts])
    out["aff_corr"] = np.array([ref.corr_cost(A, B, T) for T in Ts])
    out["aff_mi"] = np.array([ref.mi_cost(A, B, T) for T in Ts])
    hists = [ref.mi_hist(A, B, T).ravel() for T in Ts]
    idx = [np.nonzero(h)[0].astype(np.int32) for h in hists]
    out["aff_mi_hist_idx"] = np.concatenate(idx)
    out["aff_mi_hist_val"] = np.concatenate([h[i] for h, i in zip(hists, idx)]).astype(np.int32)
    out["aff_mi_hist_len"] = np.array([len(i) for i in idx], np.int32)

    q = np.random.default_rng(104).normal(0, 300, (9, 10, 11)).astype(np.float32)
    out.update(q_in=q, q_out=ref.normalize(q, 255.99))

    v = np.random.default_rng(105).integers(0, 256, (9, 10, 11)).astype(np.uint8)
    out.update(down_in=v, down_out=ref.downsample(v))
    f = np.random.default_rng(106).standard_normal((5, 6, 7, 3)).astype(np.float32)
    out.update(up_in=f, up_out=ref.upsample(f, (10, 12, 14)))

    shape = (10, 12, 14)
    Is = smooth_volume(shape, 107, background=False)
    It = smooth_volume(shape, 108)
    d = field(shape, 109, 3.0)
    Js = ref.compose(Is, d)
    out.update(cdm_Is=Is, cdm_It=It, cdm_d=d, cdm_Js=Js)
    gr, cr = ref.gradient(Js, It, "r")
    gd, cd = ref.gradient(Js, It, "d")
    out.update(cdm_grad_r=gr, cdm_cost_r=np.float32(cr), cdm_grad_d=gd, cdm_cost_d=np.float32(cd))
    nd = field(shape, 110, 1.0, 0.0)
    ps = ref.poisson(nd)
    out.update(cdm_nd=nd, cdm_poisson=ps, cdm_maxlen=np.float32(ref.max_length(ps)))
    out["cdm_accumulate"] = ref.accumulate(d, ps)
    out["cdm_smooth"] = ref.smooth(ps, d, 0.4)

    from scipy.ndimage import shift
    shape2 = (24, 20, 18)
    It2 = smooth_volume(shape2, 111, sigma=2.5)
    Is2 = shift(It2.astype(np.float32), (1.0, -1.5, 2.0), order=1).astype(np.uint8)
    ok, f2 = ref.cdm(It2, Is2)
    out.update(full_It=It2, full_Is=Is2, full_ok=np.int32(ok), full_field=f2)

    # preprocessing: otsu threshold, otsu median, chunked median, subject_image_pre
    imgs = prep_images()
    out["prep_otsu"] = np.array([ref.otsu_threshold(v) for v in imgs], np.float32)
    out["prep_otsu_median"] = np.array([ref.otsu_median(v) for v in imgs], np.float32)
    out["prep_median"] = np.array([ref.median(v) for v in imgs], np.float32)
    out["prep_subject"] = np.concatenate([ref.subject_pre(v).ravel() for v in imgs])
    big = np.random.default_rng(77).gamma(2.0, 100.0, 182 * 218 * 182).astype(np.float32)
    out["prep_median_full"] = np.float32(ref.median(big))
    q16 = np.random.default_rng(112).integers(-32768, 32767, (9, 10, 11), endpoint=True)
    q16 = q16.astype(np.int16)
    out.update(q16_in=q16, q16_out=ref.normalize(q16, 255.99))

    path = os.path.join(HERE, "reference_vectors.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    main()
