"""The nonlinear half of mm_reg composed from the oracle's functions, in the
reference's order (reg/mm_reg.hpp:285-311 calculate_linear_r and
It_match_fov, :330-346 calculate_nonlinear_r, :372-405 nonlinear_reg,
:271-284 compute_mapping_from_displacement).  Test helper only."""
import numpy as np


def oracle_mm_reg(o, It, tvs, I, svs, arg, reverse=True, match_fov=True, masked_r=True,
                  cdm_fixed=None):
    """cdm_fixed=(levels, level_iterations, smoothing, bandwidth) runs the
    fixed-schedule CDM extension in both directions instead of cdm()."""
    T = o.affine_matrix(arg, It.shape[::-1], tvs, I.shape[::-1], svs)
    J = o.resample(I, It.shape, T)
    r_lin = np.float32(o.correlation(J, It, masked_r))
    It_fov = o.match_fov(It, T, I.shape) if match_fov else np.array(It)

    def cdm(a, b):
        if cdm_fixed is None:
            return o.cdm(a, b)[1]
        lv, it, sm, bw = cdm_fixed
        return o.cdm_fixed(a, b, lv, it, 0.3, sm, bw)[0]

    t2f = cdm(It_fov, J)
    f2t = cdm(J, It_fov) if reverse else None
    to2from = o.displacement_to_mapping(t2f, T)
    from2to = o.inv_displacement_to_mapping(f2t, T, I.shape) if reverse else None
    JJ = o.compose_mapping(I, to2from)
    r_nl = np.float32(o.correlation(JJ, It_fov, masked_r))
    return dict(T=T, r=np.array([r_lin, r_nl], np.float32), J_lin=J, It_fov=It_fov, t2f=t2f,
                f2t=f2t, to2from=to2from, from2to=from2to, JJ=JJ)
