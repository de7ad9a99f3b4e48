"""The oracle against the golden vectors that tests/golden/make_golden.py
wrote from the compiled reference (tests/golden/reference_vectors.npz).
These run everywhere, including where the reference build is absent."""
import os

import numpy as np
import pytest

from conftest import prep_images

G = np.load(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.npz"))


def _bits(a):
    a = np.asarray(a)
    return a.view(np.uint32) if a.dtype == np.float32 else a


def test_affine_vectors(oracle):
    A, B, P = G["aff_A"], G["aff_B"], G["aff_params"]
    fd = np.array(A.shape[::-1], np.int32)
    td = np.array(B.shape[::-1], np.int32)
    lens = G["aff_mi_hist_len"]
    offs = np.concatenate([[0], np.cumsum(lens)])
    for k, p in enumerate(P):
        T = oracle.affine_matrix(p, fd, G["aff_fvs"], td, G["aff_tvs"])
        assert np.array_equal(_bits(T), _bits(G["aff_T"][k]))
        Ti, _ = oracle.matrix_inverse(T)
        assert np.array_equal(_bits(Ti), _bits(G["aff_Tinv"][k]))
        assert np.array_equal(oracle.resample(B, A.shape, T), G["aff_resample"][k])
        assert oracle.corr_cost(A, B, T) == G["aff_corr"][k]
        assert oracle.mi_cost(A, B, T) == G["aff_mi"][k]
        h = np.zeros(65536, np.int32)
        sl = slice(offs[k], offs[k + 1])
        h[G["aff_mi_hist_idx"][sl]] = G["aff_mi_hist_val"][sl]
        assert np.array_equal(oracle.mi_hist(A, B, T).ravel(), h)


def test_quantise_and_pyramid_vectors(oracle):
    assert np.array_equal(oracle.normalize(G["q_in"], 255.99), G["q_out"])
    assert np.array_equal(oracle.downsample(G["down_in"]), G["down_out"])
    assert np.array_equal(_bits(oracle.upsample(G["up_in"], G["up_out"].shape[:3])), _bits(G["up_out"]))


def test_cdm_step_vectors(oracle):
    Js = oracle.compose(G["cdm_Is"], G["cdm_d"])
    assert np.array_equal(Js, G["cdm_Js"])
    gr, cr = oracle.gradient(Js, G["cdm_It"], "r")
    gd, cd = oracle.gradient(Js, G["cdm_It"], "d")
    assert np.array_equal(_bits(gr), _bits(G["cdm_grad_r"]))
    assert np.array_equal(_bits(gd), _bits(G["cdm_grad_d"]))
    assert cr == pytest.approx(float(G["cdm_cost_r"]), rel=1e-6)
    assert cd == pytest.approx(float(G["cdm_cost_d"]), rel=1e-6)
    p = oracle.poisson(G["cdm_nd"])
    assert np.array_equal(_bits(p), _bits(G["cdm_poisson"]))
    assert np.float32(oracle.max_length(p)) == G["cdm_maxlen"]
    assert np.array_equal(_bits(oracle.accumulate(G["cdm_d"], p)), _bits(G["cdm_accumulate"]))
    assert np.array_equal(_bits(oracle.smooth(p, G["cdm_d"], 0.4)), _bits(G["cdm_smooth"]))


def test_whole_cdm_vector(oracle):
    ok, f, _ = oracle.cdm(G["full_It"], G["full_Is"])
    assert ok == bool(G["full_ok"])
    rms = float(np.sqrt(((f - G["full_field"]) ** 2).sum(-1).mean()))
    assert rms <= 1e-6, rms


def test_prep_vectors(oracle):
    imgs = prep_images()
    assert np.array_equal(np.array([oracle.otsu_threshold(v) for v in imgs], np.float32),
                          G["prep_otsu"])
    assert np.array_equal(np.array([oracle.otsu_median(v) for v in imgs], np.float32),
                          G["prep_otsu_median"])
    assert np.array_equal(np.array([oracle.median(v) for v in imgs], np.float32), G["prep_median"])
    assert np.array_equal(np.concatenate([oracle.subject_pre(v).ravel() for v in imgs]),
                          G["prep_subject"])
    big = np.random.default_rng(77).gamma(2.0, 100.0, 182 * 218 * 182).astype(np.float32)
    assert np.float32(oracle.median(big)) == G["prep_median_full"]
    assert np.array_equal(oracle.normalize(G["q16_in"], 255.99), G["q16_out"])


G2 = np.load(os.path.join(os.path.dirname(__file__), "golden", "cdm2d_vectors.npz"))


def test_cdm2d_vectors(oracle):
    """The 2-D CDM (cdm<out,uint8,2>) steps and one whole run against the
    vectors tests/golden/make_golden_cdm2d.py wrote from the compiled reference."""
    assert np.array_equal(oracle.downsample2(G2["down_in"]), G2["down_out"])
    assert np.array_equal(_bits(oracle.upsample2(G2["up_in"], G2["up_out"].shape[:2])),
                          _bits(G2["up_out"]))
    Js = oracle.compose2(G2["Is"], G2["d"])
    assert np.array_equal(Js, G2["Js"])
    gr, cr = oracle.gradient2(Js, G2["It"], "r")
    gd, cd = oracle.gradient2(Js, G2["It"], "d")
    assert np.array_equal(_bits(gr), _bits(G2["grad_r"]))
    assert np.array_equal(_bits(gd), _bits(G2["grad_d"]))
    assert cr == pytest.approx(float(G2["cost_r"]), rel=1e-6)
    assert cd == pytest.approx(float(G2["cost_d"]), rel=1e-6)
    p = oracle.poisson2(G2["nd"])
    assert np.array_equal(_bits(p), _bits(G2["poisson"]))
    assert np.float32(oracle.max_length2(p)) == G2["maxlen"]
    assert np.array_equal(_bits(oracle.accumulate2(G2["d"], p)), _bits(G2["accumulate"]))
    assert np.array_equal(_bits(oracle.smooth2(p, G2["d"], 0.4)), _bits(G2["smooth"]))
    ok, f, _ = oracle.cdm2(G2["full_It"], G2["full_Is"])
    assert ok == bool(G2["full_ok"])
    rms = float(np.sqrt(((f - G2["full_field"]) ** 2).sum(-1).mean()))
    assert rms <= 1e-6, rms
