// knn.hip -- photon k-nearest-neighbour gather and radiance estimate (gfx950).
//
// Same answer as PhotonMap::GetNearestNPhotons (PhotonMap.cpp:76-95,
// 224-297): the photons whose normal equals the query normal and whose
// squared distance, computed as (x*x)+(y*y)+(z*z) in float, is strictly below
// fl(maxDist*maxDist); of those, the k nearest.  The walk is the reference's
// implicit balanced tree (node = middle of its index range), but nearer child
// first and pruned with the current k-th distance instead of ENCORE's stale
// by-value radius; both only skip subtrees whose photons are at least as far
// as the splitting plane, so the selected set is the same up to photons tied
// with the k-th distance.
//
// One lane per query.  Each lane keeps a bounded max-heap of (d2, slot) in
// LDS, laid out [entry][lane] so lanes at different heap positions still hit
// different banks, and a short ring stack of pending ranges (LDS, spilling to
// HBM).  Results are written nearest first.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "rq_internal.h"

namespace rq {
namespace dev {

constexpr int kKnnWave = 64;
constexpr int kKnnStack = 8;  // LDS ring entries per lane: (lo, hi, plane d2)

__device__ __forceinline__ uint32_t canon(float f) { return f == 0.0f ? 0u : __float_as_uint(f); }

__device__ __forceinline__ uint32_t lookup_class(const uint32_t* table, uint32_t mask, float nx, float ny,
                                                 float nz) {
    if (nx != nx || ny != ny || nz != nz) return kClassNever;
    const uint32_t x = canon(nx), y = canon(ny), z = canon(nz);
    uint32_t s = normal_hash(x, y, z) & mask;
    for (uint32_t probe = 0; probe <= mask; ++probe) {
        const uint4 e = *reinterpret_cast<const uint4*>(table + 4 * s);
        if (e.w == 0xFFFFFFFFu) return kClassNever;
        if (e.x == x && e.y == y && e.z == z) return e.w;
        s = (s + 1) & mask;
    }
    return kClassNever;
}

__global__ __launch_bounds__(64) void knn_kernel(KnnLaunch L) {
    extern __shared__ uint32_t smem[];
    const int lane = threadIdx.x;
    const uint32_t k = L.k;
    float* hd = reinterpret_cast<float*>(smem) + lane;           // heap d2, entry j at hd[j*64]
    uint32_t* hs = smem + (size_t)k * kKnnWave + lane;            // heap slot
    uint32_t* st = smem + (size_t)2 * k * kKnnWave + lane;        // stack, word w of entry e at st[(e*3+w)*64]
    uint32_t* spill = L.spill + ((size_t)blockIdx.x * kKnnWave) * L.spill_depth * 3 + lane;
    for (uint32_t q = blockIdx.x * kKnnWave + lane; q < L.nq; q += gridDim.x * kKnnWave) {
    const float qx = L.qpos[3 * (size_t)q], qy = L.qpos[3 * (size_t)q + 1], qz = L.qpos[3 * (size_t)q + 2];
    const uint32_t cls = lookup_class(L.table, L.table_mask, L.qnrm[3 * (size_t)q], L.qnrm[3 * (size_t)q + 1],
                                      L.qnrm[3 * (size_t)q + 2]);
    uint32_t cnt = 0;
    float r2 = L.r2;
    if (cls != kClassNever && L.n > 0 && k > 0) {
        int sp = 0, lo_mark = 0;
        int lo = 0, hi = (int)L.n - 1;
        for (;;) {
            while (lo <= hi) {
                const int m = (lo + hi) >> 1;
                const PhotonHot p = L.hot[m];
                const float x = qx - p.x, y = qy - p.y, z = qz - p.z;
                const float d2 = (x * x + y * y) + z * z;
                if (d2 < r2 && (p.meta >> 2) == cls) {
                    if (cnt < k) {
                        // sift up
                        uint32_t i = cnt++;
                        while (i > 0) {
                            const uint32_t par = (i - 1) >> 1;
                            const float pd = hd[par * kKnnWave];
                            if (pd >= d2) break;
                            hd[i * kKnnWave] = pd;
                            hs[i * kKnnWave] = hs[par * kKnnWave];
                            i = par;
                        }
                        hd[i * kKnnWave] = d2;
                        hs[i * kKnnWave] = (uint32_t)m;
                    } else {
                        // replace the maximum, sift down
                        uint32_t i = 0;
                        for (;;) {
                            uint32_t c = 2 * i + 1;
                            if (c >= k) break;
                            float cd = hd[c * kKnnWave];
                            if (c + 1 < k) {
                                const float c2 = hd[(c + 1) * kKnnWave];
                                if (c2 > cd) {
                                    cd = c2;
                                    ++c;
                                }
                            }
                            if (cd <= d2) break;
                            hd[i * kKnnWave] = cd;
                            hs[i * kKnnWave] = hs[c * kKnnWave];
                            i = c;
                        }
                        hd[i * kKnnWave] = d2;
                        hs[i * kKnnWave] = (uint32_t)m;
                    }
                    if (cnt == k) r2 = hd[0];
                }
                const uint32_t axis = p.meta & 3u;
                const float pa = axis == 0 ? p.x : (axis == 1 ? p.y : p.z);
                const float qa = axis == 0 ? qx : (axis == 1 ? qy : qz);
                const float dp = qa - pa;
                const float dp2 = dp * dp;
                // Nearer side first: the reference's one-sided choice
                // (PhotonMap.cpp:276-289) goes right when dp > 0.
                int nlo, nhi, flo, fhi;
                if (dp > 0.0f) {
                    nlo = m + 1; nhi = hi; flo = lo; fhi = m - 1;
                } else {
                    nlo = lo; nhi = m - 1; flo = m + 1; fhi = hi;
                }
                if (dp2 < r2 && flo <= fhi) {
                    // push (flo, fhi, dp2)
                    if (sp - lo_mark == kKnnStack) {
                        const int s = lo_mark & (kKnnStack - 1);
                        for (int w = 0; w < 3; ++w)
                            spill[((size_t)lo_mark * 3 + w) * kKnnWave] = st[(s * 3 + w) * kKnnWave];
                        ++lo_mark;
                    }
                    const int s = sp & (kKnnStack - 1);
                    st[(s * 3 + 0) * kKnnWave] = (uint32_t)flo;
                    st[(s * 3 + 1) * kKnnWave] = (uint32_t)fhi;
                    st[(s * 3 + 2) * kKnnWave] = __float_as_uint(dp2);
                    ++sp;
                }
                lo = nlo;
                hi = nhi;
            }
            // pop the next range that can still hold a closer photon
            bool found = false;
            while (sp > 0) {
                --sp;
                uint32_t e0, e1, e2;
                if (sp < lo_mark) {
                    lo_mark = sp;
                    e0 = spill[((size_t)sp * 3 + 0) * kKnnWave];
                    e1 = spill[((size_t)sp * 3 + 1) * kKnnWave];
                    e2 = spill[((size_t)sp * 3 + 2) * kKnnWave];
                } else {
                    const int s = sp & (kKnnStack - 1);
                    e0 = st[(s * 3 + 0) * kKnnWave];
                    e1 = st[(s * 3 + 1) * kKnnWave];
                    e2 = st[(s * 3 + 2) * kKnnWave];
                }
                if (__uint_as_float(e2) < r2) {
                    lo = (int)e0;
                    hi = (int)e1;
                    found = true;
                    break;
                }
            }
            if (!found) break;
        }
    }
    // Heap sort into ascending order: the maximum goes to the end each round.
    const size_t base = (size_t)q * k;
    double R = 0.0, G = 0.0, B = 0.0;
    const double rmax = cnt ? (double)hd[0] : 0.0;
    for (uint32_t end = cnt; end > 0; --end) {
        const float topd = hd[0];
        const uint32_t tops = hs[0];
        if (L.out_idx) {
            L.out_idx[base + end - 1] = L.perm[tops];
            L.out_d2[base + end - 1] = topd;
        }
        if (L.out_rgb) {
            const float4 pw = *reinterpret_cast<const float4*>(L.power + tops);
            const double w = L.caustic ? (rmax - (double)topd) / (double)topd : 1.0;
            R += (double)pw.x * w;
            G += (double)pw.y * w;
            B += (double)pw.z * w;
        }
        // move the last leaf to the root and sift down over [0, end-1)
        const uint32_t n2 = end - 1;
        if (n2 == 0) break;
        const float vd = hd[n2 * kKnnWave];
        const uint32_t vs = hs[n2 * kKnnWave];
        uint32_t i = 0;
        for (;;) {
            uint32_t c = 2 * i + 1;
            if (c >= n2) break;
            float cd = hd[c * kKnnWave];
            if (c + 1 < n2) {
                const float c2 = hd[(c + 1) * kKnnWave];
                if (c2 > cd) {
                    cd = c2;
                    ++c;
                }
            }
            if (cd <= vd) break;
            hd[i * kKnnWave] = cd;
            hs[i * kKnnWave] = hs[c * kKnnWave];
            i = c;
        }
        hd[i * kKnnWave] = vd;
        hs[i * kKnnWave] = vs;
    }
    if (L.out_idx) {
        for (uint32_t j = cnt; j < k; ++j) {
            L.out_idx[base + j] = 0xFFFFFFFFu;
            L.out_d2[base + j] = __int_as_float(0x7f800000);
        }
    }
    if (L.out_cnt) L.out_cnt[q] = cnt;
    if (L.out_rgb) {
        float r = 0.0f, g = 0.0f, b = 0.0f;
        if (cnt) {
            r = (float)R;
            g = (float)G;
            b = (float)B;
            if (L.caustic) {
                // N / (PI r^2) with ENCORE's PI = 3.14159f and r the largest
                // squared distance (PhotonMap.cpp:135, 157).
                const float dens = (float)((double)k / ((double)3.14159f * rmax * rmax));
                r *= dens;
                g *= dens;
                b *= dens;
            }
        }
        L.out_rgb[3 * (size_t)q] = r;
        L.out_rgb[3 * (size_t)q + 1] = g;
        L.out_rgb[3 * (size_t)q + 2] = b;
    }
    }  // queries
}

}  // namespace dev

uint32_t knn_spill_words_per_lane(uint32_t n) {
    uint32_t d = 2;
    while (n) {
        ++d;
        n >>= 1;
    }
    return d;  // stack entries per lane >= tree depth
}

size_t knn_lds_bytes(uint32_t k) { return ((size_t)2 * k + 3 * dev::kKnnStack) * dev::kKnnWave * sizeof(uint32_t); }

int knn_grid(uint32_t nq, int num_cus, uint32_t k) {
    int per_cu = 0;
    RQ_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)dev::knn_kernel, dev::kKnnWave,
                                                        knn_lds_bytes(k)));
    const int64_t need = ((int64_t)nq + dev::kKnnWave - 1) / dev::kKnnWave;
    return (int)std::max<int64_t>(1, std::min<int64_t>(need, (int64_t)std::max(1, per_cu) * num_cus));
}

void launch_knn(const KnnLaunch& L, int grid, hipStream_t s) {
    if (L.nq == 0) return;
    const size_t lds = knn_lds_bytes(L.k);
    hipLaunchKernelGGL(dev::knn_kernel, dim3(grid), dim3(dev::kKnnWave), lds, s, L);
    RQ_HIP(hipGetLastError());
}

}  // namespace rq
