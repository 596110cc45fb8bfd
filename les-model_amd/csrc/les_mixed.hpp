// les_mixed.hpp -- FP32 workspaces of the mixed-precision PCG.
#pragma once

#include <hip/hip_runtime.h>

namespace les {

struct MixedWork {
    float *r = nullptr, *p = nullptr, *p2 = nullptr, *q = nullptr, *z = nullptr, *za = nullptr,
          *zb = nullptr, *pc = nullptr, *pc2 = nullptr, *x = nullptr;
};

inline void free_mixed(MixedWork& w)
{
    for (float* p : {w.r, w.p, w.p2, w.q, w.z, w.za, w.zb, w.pc, w.pc2, w.x})
        if (p) hipFree(p);
    w = MixedWork{};
}

}  // namespace les
