// layout.hip -- reorder between the reference's host layout (t-fastest
// lexicographic AoS, source/MPILattice/LocalLayout.h:149-155,
// source/LatticeDefinitions.h:86-114) and the device even-odd SoA layout.
#include "lym_kernels.h"

namespace lym {

__device__ __forceinline__ void lex_coords(const Geom& g, long lex, int c[4]) {
    c[3] = (int)(lex % g.L[3]);
    long r = lex / g.L[3];
    c[2] = (int)(r % g.L[2]);
    r /= g.L[2];
    c[1] = (int)(r % g.L[1]);
    c[0] = (int)(r / g.L[1]);
}

__global__ void k_spinor_to_eo(Geom g, const zd* __restrict__ aos, zd* __restrict__ eo) {
    long lex = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (lex >= ((long)g.Vh << g.tsh)) return;
    int c[4];
    lex_coords(g, lex, c);
    int p = site_parity(g, c);
    int idx = eo_index(g, c[0], c[1], c[2], c[3]);
    zd* o = eo + (size_t)p * 12 * g.stride + idx;
#pragma unroll
    for (int k = 0; k < 12; ++k) o[(size_t)k * g.stride] = aos[lex * 12 + k];
}

__global__ void k_spinor_from_eo(Geom g, const zd* __restrict__ eo, zd* __restrict__ aos) {
    long lex = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (lex >= ((long)g.Vh << g.tsh)) return;
    int c[4];
    lex_coords(g, lex, c);
    int p = site_parity(g, c);
    int idx = eo_index(g, c[0], c[1], c[2], c[3]);
    const zd* s = eo + (size_t)p * 12 * g.stride + idx;
#pragma unroll
    for (int k = 0; k < 12; ++k) aos[lex * 12 + k] = s[(size_t)k * g.stride];
}

// links: host element (i,j) at i + 3j (column-major); device component 3i + j
template <class E>
__global__ void k_links_to_eo(Geom g, const E* __restrict__ aos, int ncomp, int unfold_bc, E* __restrict__ eo) {
    long lex = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (lex >= ((long)g.Vh << g.tsh)) return;
    int c[4];
    lex_coords(g, lex, c);
    int p = site_parity(g, c);
    int idx = eo_index(g, c[0], c[1], c[2], c[3]);
    for (int mu = 0; mu < 4; ++mu) {
        const E* s = aos + (lex * 4 + mu) * 9;
        E* o = eo + (size_t)((p * 4 + mu) * ncomp) * g.stride + idx;
        bool flip = unfold_bc && g.bc_t < 0 && mu == 3 && c[3] == g.L[3] - 1;
        for (int k = 0; k < ncomp; ++k) {
            int i = k / 3, j = k % 3;
            E v = s[i + 3 * j];
            if (flip) v = -v;
            o[(size_t)k * g.stride] = v;
        }
    }
}

void launch_spinor_to_eo(const Geom& g, const zd* aos, zd* eo, hipStream_t st) {
    long V = (long)g.Vh << g.tsh;
    hipLaunchKernelGGL(k_spinor_to_eo, dim3((V + 255) / 256), dim3(256), 0, st, g, aos, eo);
}

void launch_spinor_from_eo(const Geom& g, const zd* eo, zd* aos, hipStream_t st) {
    long V = (long)g.Vh << g.tsh;
    hipLaunchKernelGGL(k_spinor_from_eo, dim3((V + 255) / 256), dim3(256), 0, st, g, eo, aos);
}

void launch_links_to_eo(const Geom& g, const void* links_aos, int real_links, int ncomp, int unfold_bc, void* eo,
                        hipStream_t st) {
    long V = (long)g.Vh << g.tsh;
    if (real_links)
        hipLaunchKernelGGL(k_links_to_eo<double>, dim3((V + 255) / 256), dim3(256), 0, st, g,
                           (const double*)links_aos, ncomp, unfold_bc, (double*)eo);
    else
        hipLaunchKernelGGL(k_links_to_eo<zd>, dim3((V + 255) / 256), dim3(256), 0, st, g, (const zd*)links_aos,
                           ncomp, unfold_bc, (zd*)eo);
}

}  // namespace lym
