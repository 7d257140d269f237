// Sign encoding of the contact pair list (_collisonPairs, GIPC.cuh:49-51, as
// read by TraditionalMAS32Preconditioner.cu:1550-1960): four ints per pair,
// the signs of x, y, z, w select the pair type.  Shared by the host analysis
// and the device hierarchy kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace masx
{
// Writes the pair's vertex ids to v and returns how many there are (2, 3 or 4).
__host__ __device__ inline int decode_pair_code(const int* p, int v[4])
{
    const int x = p[0], y = p[1], z = p[2], w = p[3];
    if(x >= 0)
    {
        v[0] = x; v[1] = y; v[2] = z; v[3] = w < 0 ? -w - 1 : w;
        return 4;
    }
    v[0] = -x - 1;
    if(z < 0)
    {
        if(y < 0)
        {
            v[1] = -y - 1; v[2] = -z - 1; v[3] = -w - 1;
            return 4;
        }
        v[1] = y;
        return 2;
    }
    if(w < 0)
    {
        if(y < 0)
        {
            v[1] = -y - 1; v[2] = z; v[3] = -w - 1;
            return 4;
        }
        v[1] = y; v[2] = z;
        return 3;
    }
    v[1] = y; v[2] = z; v[3] = w;
    return 4;
}
}  // namespace masx
