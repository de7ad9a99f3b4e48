// cdm_control.hpp -- the host-side control of the CDM loop shared by the
// 3-D (cdm.hip) and 2-D (cdm2d.hip) paths.  Plain C++ (no HIP).
#pragma once

#include <cstdint>
#include <vector>

namespace tiplhip {

// cdm_improved (reg/cdm.hpp:19-34) with linear_regression over (index, cost)
// (numerical/statistics.hpp:694-713): false when the cost trend rises;
// drops the oldest entry once more than 7 are held.
inline bool cdm_improved(std::vector<float>& cost)
{
    const size_t sz = cost.size();
    if (sz > 5) {
        int64_t sx = 0, sxx = 0;
        double sy = 0.0, syy = 0.0, sxy = 0.0;
        for (size_t i = 0; i < sz; ++i) {
            sx += int64_t(i);
            sxx += int64_t(i) * int64_t(i);
            sy += double(cost[i]);
            syy += double(cost[i]) * double(cost[i]);
            sxy += double(float(int(i)) * cost[i]);
        }
        const double dn = double(sz);
        const double mean_x = double(sx) / dn, mean_y = sy / dn;
        const double x_var = double(sxx) / dn - mean_x * mean_x;
        const double y_var = syy / dn - mean_y * mean_y;
        float a = 0.0f;
        if (!(x_var == 0.0 || y_var == 0.0)) a = float((sxy / dn - mean_x * mean_y) / x_var);
        if (a > 0.0f) return false;
        if (sz > 7) cost.erase(cost.begin());
    }
    return true;
}

}  // namespace tiplhip
