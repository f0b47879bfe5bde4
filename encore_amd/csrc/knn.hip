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
// Three kernels give that answer:
//  - knn_kernel (k < 32, or no cell grid): one lane per query, a bounded
//    max-heap of (d2, slot) in LDS laid out [entry][lane] so lanes at
//    different heap positions still hit different banks, and a short ring
//    stack of pending ranges (LDS, spilling to HBM);
//  - knn_cell_kernel (k >= 32): one query per wave over the photon cell grid
//    of photon_grid.hip, lists and/or radiance;
//  - knn_wave_kernel (k >= 32, radiance only, when no grid suits the radius):
//    one query per wave walking the tree.
// Lists are written nearest first.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdlib>

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
    for (uint32_t qi = blockIdx.x * kKnnWave + lane; qi < L.nq; qi += gridDim.x * kKnnWave) {
    const uint32_t q = L.order ? L.order[qi] : qi;  // neighbouring lanes take neighbouring queries
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

// ----- large k, radiance only: one query per wave -------------------------
// For k in the hundreds a per-lane heap needs k x 8 bytes of LDS per lane,
// which leaves a CU two or three waves.  Here the 64 lanes of a wave serve
// one query together:
//  - the tree is walked 64 ranges per step from a stack of pending ranges
//    (global memory, one region per wave, popped 64 at a time from the top);
//  - photons that pass the class and distance tests go to a candidate buffer
//    in LDS;
//  - once the buffer holds 2k candidates, the k-th smallest squared distance
//    is found by a radix select over the float bits, the buffer keeps the k
//    smallest and that distance becomes the strict bound for new candidates
//    and for pruning, as the heap's maximum does in knn_kernel;
//  - at the end the k smallest are kept and their power summed in double.
// Ties at the k-th distance keep the candidates found first.  The stack keeps
// its entries in non-decreasing tree depth (children are pushed in the order
// of their parents), so a depth level never holds more than 128 entries and
// the stack at most 128 x (depth + 1).
constexpr int kWaveCand = 1024;

struct WaveStack {
    uint32_t* lo;
    uint32_t* hi;
    uint32_t* dp;
};

// Stack entries written by this wave's earlier steps: loads that bypass the
// CU's cache, after the stores were made to complete.
__device__ __forceinline__ uint32_t stack_load(uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
    return (uint32_t)__popcll(m & ((1ull << __lane_id()) - 1ull));
}

// Radix select of the k smallest keys (k = the 1-based rank `need` on entry)
// among s_key[0, n), 8 bits per pass from the top: returns a threshold t and
// leaves in `need` how many entries equal to t belong to the k smallest, so
// that they are the keys below t and the first `need` keys equal to t.  t is
// the k-th smallest key, or, once every key of the k-th key's bucket belongs
// to the k smallest, the end of that bucket with need = 0 (the remaining
// passes would change nothing).  Either way no key of the k smallest exceeds
// t, so t bounds the search radius from above.
__device__ uint32_t select_kth(const uint32_t* s_key, uint32_t n, uint32_t& need, uint32_t* s_hist) {
    const uint32_t lane = __lane_id();
    uint32_t prefix = 0, mask = 0;
    for (int shift = 24; shift >= 0; shift -= 8) {
        for (uint32_t b = lane; b < 256; b += 64) s_hist[b] = 0;
        __builtin_amdgcn_wave_barrier();
        for (uint32_t i = lane; i < n; i += 64) {
            const uint32_t key = s_key[i];
            if ((key & mask) == prefix) atomicAdd(&s_hist[(key >> shift) & 255u], 1u);
        }
        __builtin_amdgcn_wave_barrier();
        // each lane sums 4 bins; a scan over lanes finds the digit
        uint32_t c[4], loc = 0;
        for (int j = 0; j < 4; ++j) {
            c[j] = s_hist[4 * lane + j];
            loc += c[j];
        }
        uint32_t incl = loc;
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t v = __shfl_up(incl, off, 64);
            if (lane >= (uint32_t)off) incl += v;
        }
        const uint32_t excl = incl - loc;
        const bool here = excl < need && incl >= need;
        const uint64_t hm = __ballot(here);
        if (hm == 0) break;  // need exceeds the keys left: cannot happen for need <= n
        const uint32_t src = (uint32_t)__ffsll((unsigned long long)hm) - 1;
        uint32_t digit = 0, before = 0, in_bucket = 0;
        if (here) {
            uint32_t run = excl;
            for (int j = 0; j < 4; ++j) {
                if (run + c[j] >= need) {
                    digit = 4 * lane + j;
                    before = run;
                    in_bucket = c[j];
                    break;
                }
                run += c[j];
            }
        }
        digit = __shfl(digit, src, 64);
        before = __shfl(before, src, 64);
        in_bucket = __shfl(in_bucket, src, 64);
        need -= before;
        prefix |= digit << shift;
        mask |= 255u << shift;
        __builtin_amdgcn_wave_barrier();
        // The caller reads t as a float bound, so end may be taken only while
        // it is a float in [0, +inf]: a bucket of keys at or above 2^127
        // would end at 0x80000000, which reads as -0.  (The callers pass more
        // than k keys, all finite squared distances, so the bucket of the
        // k-th key holds keys above it and never passes this test at that
        // end; the bound keeps it so for any input.)
        const uint64_t end = (uint64_t)prefix + (1ull << shift);
        if (need == in_bucket && end <= 0x7F800000ull) {
            need = 0;
            return (uint32_t)end;
        }
    }
    return prefix;
}

// Keeps, in buffer order, the keys below t and the first `need_eq` equal to t.
__device__ uint32_t compact_kth(uint32_t* s_key, uint32_t* s_slot, uint32_t n, uint32_t t, uint32_t need_eq) {
    uint32_t out = 0, eq_seen = 0;
    for (uint32_t i0 = 0; i0 < n; i0 += 64) {
        const uint32_t i = i0 + __lane_id();
        const bool valid = i < n;
        const uint32_t key = valid ? s_key[i] : 0xFFFFFFFFu;
        const uint32_t slot = valid ? s_slot[i] : 0u;
        const bool lt = valid && key < t, eq = valid && key == t;
        const uint64_t em = __ballot(eq);
        const bool keep = lt || (eq && eq_seen + lanes_below(em) < need_eq);
        const uint64_t km = __ballot(keep);
        __builtin_amdgcn_wave_barrier();
        if (keep) {
            const uint32_t pos = out + lanes_below(km);
            s_key[pos] = key;
            s_slot[pos] = slot;
        }
        __builtin_amdgcn_wave_barrier();
        out += (uint32_t)__popcll(km);
        eq_seen += (uint32_t)__popcll(em);
    }
    return out;
}

// The largest of the n keys left after a compaction: the k-th smallest
// squared distance so far, the strict bound for later candidates (ENCORE
// takes the heap's maximum once the heap is full, PhotonMap.cpp:255, 268).
// The selection's threshold can be the end of a radix bucket, up to 4x that.
__device__ __forceinline__ uint32_t kept_max(const uint32_t* s_key, uint32_t n) {
    uint32_t m = 0;
    for (uint32_t i = __lane_id(); i < n; i += 64) m = max(m, s_key[i]);
    for (int off = 32; off > 0; off >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, off, 64));
    return m;
}

// The value held by lane (lane ^ J) of the wave: DPP quad permutes for
// J = 1, 2, LDS swizzles within 32 lanes for J = 4 .. 16, a permute for 32.
template <uint32_t J>
__device__ __forceinline__ uint32_t lane_xor(uint32_t x) {
    if constexpr (J == 1) return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0xB1, 0xF, 0xF, false);
    else if constexpr (J == 2) return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x4E, 0xF, 0xF, false);
    else if constexpr (J <= 16) return (uint32_t)__builtin_amdgcn_ds_swizzle((int)x, 0x1F | (J << 10));
    else return (uint32_t)__builtin_amdgcn_ds_bpermute((int)((__lane_id() ^ J) << 2), (int)x);
}

template <uint32_t J>
__device__ __forceinline__ uint64_t lane_xor(uint64_t x) {
    return (uint64_t)lane_xor<J>((uint32_t)x) | ((uint64_t)lane_xor<J>((uint32_t)(x >> 32)) << 32);
}

template <uint32_t J>
__device__ __forceinline__ double lane_xor(double x) {
    return __longlong_as_double((long long)lane_xor<J>((uint64_t)__double_as_longlong(x)));
}

// The lanes at which a compare-exchange step of a bitonic sort keeps the
// smaller of the pair: element e (lane e of the first register, lane e - 64
// of the second, BASE = 64) pairs with element e ^ J within a run of SIZE
// elements that is sorted ascending when e & SIZE is 0 (always, for SIZE =
// 128), and keeps the smaller when its own bit J is clear in an ascending run
// or set in a descending one.  A compile-time constant: the step selects with
// it as a lane mask instead of testing lane bits.
template <uint32_t SIZE, uint32_t J, uint32_t BASE>
constexpr uint64_t bitonic_min_lanes() {
    uint64_t m = 0;
    for (uint32_t l = 0; l < 64; ++l) {
        const uint32_t e = l + BASE;
        const bool up = SIZE >= 128 || (e & SIZE) == 0;
        if (((e & J) == 0) == up) m |= 1ull << l;
    }
    return m;
}

// One compare-exchange step: a lane keeps its own value when (own < other)
// agrees with keeping the smaller, else the other's (equal values: either).
template <uint32_t SIZE, uint32_t J, uint32_t BASE, typename T>
__device__ __forceinline__ T bitonic_step(T x) {
    const T p = lane_xor<J>(x);
    const uint64_t keep_own = ~(__ballot(x < p) ^ bitonic_min_lanes<SIZE, J, BASE>());
    return __builtin_amdgcn_inverse_ballot_w64(keep_own) ? x : p;
}

template <uint32_t SIZE, uint32_t J, typename T>
__device__ __forceinline__ void bitonic_merge(T& x0, T& x1, bool two) {
    if constexpr (J == 64) {
        const T lo = min(x0, x1), hi = max(x0, x1);
        x0 = lo;
        x1 = hi;
    } else {
        x0 = bitonic_step<SIZE, J, 0>(x0);
        if (two) x1 = bitonic_step<SIZE, J, 64>(x1);
    }
    if constexpr (J > 1) bitonic_merge<SIZE, J / 2>(x0, x1, two);
}

// Sorts 64 keys (one per lane) ascending, or 128 (x0 then x1) when `two`
// (wave-uniform).  Every lane of the wave takes part.
template <uint32_t SIZE = 2, typename T>
__device__ __forceinline__ void bitonic_sort(T& x0, T& x1, bool two) {
    if constexpr (SIZE <= 128) {
        if (SIZE == 128 && !two) return;
        bitonic_merge<SIZE, SIZE / 2>(x0, x1, two);
        bitonic_sort<SIZE * 2>(x0, x1, two);
    }
}

// Butterfly sum over the wave (lane ^ 32 first): every lane gets the same
// total, in the same order of additions.
__device__ __forceinline__ double wave_sum(double v) {
    v += lane_xor<32>(v);
    v += lane_xor<16>(v);
    v += lane_xor<8>(v);
    v += lane_xor<4>(v);
    v += lane_xor<2>(v);
    v += lane_xor<1>(v);
    return v;
}

// Lane 0 writes the count and the radiance of the wave's query from the
// wave's sums: the caustic filter's density factor applied
// (PhotonMap.cpp:156-161), NaN when the walk overflowed.
__device__ __forceinline__ void wave_radiance_out(const KnnLaunch& L, uint32_t q, uint32_t cand, bool overflow,
                                                  double rmax, double R, double G, double B) {
    if (__lane_id() != 0) return;
    float r = 0.0f, g = 0.0f, b = 0.0f;
    if (cand) {
        r = (float)R;
        g = (float)G;
        b = (float)B;
        if (L.caustic) {
            const float dens = (float)((double)L.k / ((double)3.14159f * rmax * rmax));
            r *= dens;
            g *= dens;
            b *= dens;
        }
    }
    if (overflow) r = g = b = __int_as_float(0x7fc00000);
    if (L.out_cnt) L.out_cnt[q] = cand;
    L.out_rgb[3 * (size_t)q] = r;
    L.out_rgb[3 * (size_t)q + 1] = g;
    L.out_rgb[3 * (size_t)q + 2] = b;
}

// The end of a one-query-per-wave gather: keeps the k nearest of the `cand`
// candidates in LDS (ties at the k-th distance: those found first) and
// writes the count, the list nearest first when asked, and the radiance when
// asked (GetRadianceEstimate, PhotonMap.cpp:121-164), NaN when the walk
// overflowed.
__device__ void wave_finish(const KnnLaunch& L, uint32_t q, uint32_t cand, bool overflow, uint32_t* s_key,
                            uint32_t* s_slot, uint32_t* s_hist) {
    const uint32_t lane = __lane_id();
    const uint32_t k = L.k;
    if (cand > (k <= 64u ? 128u : k)) {
        uint32_t need = k;
        const uint32_t t = select_kth(s_key, cand, need, s_hist);
        cand = compact_kth(s_key, s_slot, cand, t, need);
    }
    if (k <= 64u) {
        // At most 128 candidates: sort (squared distance, buffer position)
        // and keep the first k -- the same photons as the selection above
        // (ties at the k-th distance: those found first), already in list
        // order.  Sorted position `lane` holds the lane's photon.
        uint64_t x0 = lane < cand ? ((uint64_t)s_key[lane] << 7) | lane : ~0ull;
        uint64_t x1 = lane + 64u < cand ? ((uint64_t)s_key[lane + 64u] << 7) | (lane + 64u) : ~0ull;
        bitonic_sort(x0, x1, cand > 64u);
        const uint32_t kept = min(cand, k);
        const uint32_t key = (uint32_t)(x0 >> 7);
        const uint32_t slot = lane < kept ? s_slot[(uint32_t)x0 & 127u] : 0u;
        if (L.out_idx) {
            const size_t at = (size_t)q * k + lane;
            if (lane < kept) {
                L.out_idx[at] = L.perm[slot];
                L.out_d2[at] = __uint_as_float(key);
            } else if (lane < k) {
                L.out_idx[at] = 0xFFFFFFFFu;
                L.out_d2[at] = __int_as_float(0x7f800000);
            }
        }
        if (!L.out_rgb) {
            if (lane == 0 && L.out_cnt) L.out_cnt[q] = kept;
            return;
        }
        // radiance over the kept set (GetRadianceEstimate, PhotonMap.cpp:121-164)
        const double rmax = kept ? (double)__uint_as_float(__builtin_amdgcn_readlane(key, kept - 1u)) : 0.0;
        double R = 0.0, G = 0.0, B = 0.0;
        if (lane < kept) {
            const float4 pw = *reinterpret_cast<const float4*>(L.power + slot);
            const double dd = (double)__uint_as_float(key);
            const double wgt = L.caustic ? (rmax - dd) / dd : 1.0;
            R = (double)pw.x * wgt;
            G = (double)pw.y * wgt;
            B = (double)pw.z * wgt;
        }
        wave_radiance_out(L, q, kept, overflow, rmax, wave_sum(R), wave_sum(G), wave_sum(B));
        return;
    }
    if (L.out_idx) {
        // the list, nearest first: the rank of each kept photon by (d2, buffer position)
        const size_t base = (size_t)q * k;
        for (uint32_t i = lane; i < cand; i += 64) {
            const uint32_t key = s_key[i];
            uint32_t rank = 0;
            for (uint32_t j = 0; j < cand; ++j) {
                const uint32_t o = s_key[j];
                rank += (o < key || (o == key && j < i)) ? 1u : 0u;
            }
            L.out_idx[base + rank] = L.perm[s_slot[i]];
            L.out_d2[base + rank] = __uint_as_float(key);
        }
        for (uint32_t j = cand + lane; j < k; j += 64) {
            L.out_idx[base + j] = 0xFFFFFFFFu;
            L.out_d2[base + j] = __int_as_float(0x7f800000);
        }
    }
    if (!L.out_rgb) {
        if (lane == 0 && L.out_cnt) L.out_cnt[q] = cand;
        return;
    }
    // radiance over the kept set (GetRadianceEstimate, PhotonMap.cpp:121-164)
    uint32_t mx = 0;
    for (uint32_t i = lane; i < cand; i += 64) mx = max(mx, s_key[i]);
    for (int off = 32; off > 0; off >>= 1) mx = max(mx, (uint32_t)__shfl_xor((int)mx, off, 64));
    const double rmax = cand ? (double)__uint_as_float(mx) : 0.0;
    double R = 0.0, G = 0.0, B = 0.0;
    for (uint32_t i = lane; i < cand; i += 64) {
        const float4 pw = *reinterpret_cast<const float4*>(L.power + s_slot[i]);
        const double dd = (double)__uint_as_float(s_key[i]);
        const double wgt = L.caustic ? (rmax - dd) / dd : 1.0;
        R += (double)pw.x * wgt;
        G += (double)pw.y * wgt;
        B += (double)pw.z * wgt;
    }
    wave_radiance_out(L, q, cand, overflow, rmax, wave_sum(R), wave_sum(G), wave_sum(B));
}

__global__ __launch_bounds__(64) void knn_wave_kernel(KnnLaunch L, uint32_t stack_cap) {
    __shared__ uint32_t s_key[kWaveCand];   // squared distance bits (positive floats order as integers)
    __shared__ uint32_t s_slot[kWaveCand];
    __shared__ uint32_t s_hist[256];
    const uint32_t lane = __lane_id();
    const uint32_t k = L.k;
    const uint32_t compact_at = min((uint32_t)kWaveCand - 64u, max(2u * k, k + 64u));
    WaveStack st;
    st.lo = L.spill + (size_t)blockIdx.x * 3 * stack_cap;
    st.hi = st.lo + stack_cap;
    st.dp = st.hi + stack_cap;
    for (uint32_t qi = blockIdx.x; qi < L.nq; qi += gridDim.x) {
        const uint32_t q = L.order ? L.order[qi] : qi;
        const float qx = L.qpos[3 * (size_t)q], qy = L.qpos[3 * (size_t)q + 1], qz = L.qpos[3 * (size_t)q + 2];
        const uint32_t cls = lookup_class(L.table, L.table_mask, L.qnrm[3 * (size_t)q], L.qnrm[3 * (size_t)q + 1],
                                          L.qnrm[3 * (size_t)q + 2]);
        uint32_t cand = 0;
        bool overflow = false;
        if (cls != kClassNever && L.n > 0) {
            float r2 = L.r2;
            if (lane == 0) {
                st.lo[0] = 0;
                st.hi[0] = L.n - 1;
                st.dp[0] = 0;
            }
            uint32_t sp = 1;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            while (sp > 0) {
                const uint32_t take = min(sp, 64u);
                const uint32_t base = sp - take;
                sp = base;
                bool act = lane < take;
                uint32_t lo = 0, hi = 0, dpb = 0;
                if (act) {
                    lo = stack_load(st.lo + base + lane);
                    hi = stack_load(st.hi + base + lane);
                    dpb = stack_load(st.dp + base + lane);
                    act = __uint_as_float(dpb) < r2;
                }
                bool is_c = false, push_n = false, push_f = false;
                float d2 = 0.0f, dp2 = 0.0f;
                int im = 0, nlo = 0, nhi = -1, flo = 0, fhi = -1;
                if (act) {
                    const int ilo = (int)lo, ihi = (int)hi;
                    im = (ilo + ihi) >> 1;
                    const PhotonHot p = L.hot[im];
                    const float x = qx - p.x, y = qy - p.y, z = qz - p.z;
                    d2 = (x * x + y * y) + z * z;
                    is_c = d2 < r2 && (p.meta >> 2) == cls;
                    const uint32_t axis = p.meta & 3u;
                    const float pa = axis == 0 ? p.x : (axis == 1 ? p.y : p.z);
                    const float qa = axis == 0 ? qx : (axis == 1 ? qy : qz);
                    const float dp = qa - pa;
                    dp2 = dp * dp;
                    if (dp > 0.0f) {  // the query's side (PhotonMap.cpp:276-289)
                        nlo = im + 1; nhi = ihi; flo = ilo; fhi = im - 1;
                    } else {
                        nlo = ilo; nhi = im - 1; flo = im + 1; fhi = ihi;
                    }
                    push_n = nlo <= nhi;
                    push_f = dp2 < r2 && flo <= fhi;
                }
                const uint32_t m = (uint32_t)im;
                // candidates, in lane order
                const uint64_t cm = __ballot(is_c);
                if (is_c) {
                    const uint32_t pos = cand + lanes_below(cm);
                    s_key[pos] = __float_as_uint(d2);
                    s_slot[pos] = m;
                }
                cand += (uint32_t)__popcll(cm);
                // children, in the order of their parents
                const uint32_t np = (push_n ? 1u : 0u) + (push_f ? 1u : 0u);
                uint32_t incl = np;
                for (int off = 1; off < 64; off <<= 1) {
                    const uint32_t v = __shfl_up(incl, off, 64);
                    if (lane >= (uint32_t)off) incl += v;
                }
                const uint32_t total = __shfl(incl, 63, 64);
                if (sp + total > stack_cap) {
                    overflow = true;  // cannot happen within the depth bound; reported as NaN
                    break;
                }
                uint32_t w = sp + incl - np;
                if (push_n) {
                    st.lo[w] = (uint32_t)nlo;
                    st.hi[w] = (uint32_t)nhi;
                    st.dp[w] = 0u;
                    ++w;
                }
                if (push_f) {
                    st.lo[w] = (uint32_t)flo;
                    st.hi[w] = (uint32_t)fhi;
                    st.dp[w] = __float_as_uint(dp2);
                }
                sp += total;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                __builtin_amdgcn_wave_barrier();
                if (cand >= compact_at) {
                    uint32_t need = k;
                    const uint32_t t = select_kth(s_key, cand, need, s_hist);
                    cand = compact_kth(s_key, s_slot, cand, t, need);
                    r2 = __uint_as_float(kept_max(s_key, cand));
                }
            }
        }
        wave_finish(L, q, cand, overflow, s_key, s_slot, s_hist);
        __builtin_amdgcn_wave_barrier();
    }
}

// ----- k >= 32, lists and/or radiance, from the photon cell grid ----------
// One query per wave, as knn_wave_kernel, but the candidates come from the
// cells within reach of the query instead of a tree walk:
//  - lane j takes cell (j & 3, (j >> 2) & 3, j >> 4) of the 4 x 4 x 4 block
//    that covers [q - reach, q + reach] (the cell edge is the reach, so no
//    axis spans more than 4 cells), drops it when it is outside the grid or
//    farther than the radius, and looks up its photon range;
//  - where the kept cells hold several 64-photon blocks, the blocks whose
//    boxes lie within the radius replace the cells (at most 128 ranges);
//  - the ranges are ordered nearest first; blocks are read one per 64-lane
//    chunk, cells 64 photons at a time as one concatenated range;
//  - the candidates are cut to the k nearest whenever k + 64 have gathered,
//    and the k-th distance then bounds later candidates; once it has shrunk
//    below the distance of the next range, that range and every later one
//    are skipped;
//  - the final selection and the outputs are wave_finish's.
// `near` is a lower bound on the float d2 of every photon of a cell: the
// distance to the cell's box, widened for the rounding of cell assignment,
// squared and lowered by 2^-20 for the rounding of the float d2.
__device__ __forceinline__ uint2 cell_lookup(const PhotonGridView& G, uint64_t key) {
    uint32_t h = cell_hash(key, G.cell_mask);
    for (uint32_t probe = 0; probe <= G.cell_mask; ++probe) {
        const uint4 e = G.cells[h];
        const uint64_t ek = (uint64_t)e.y << 32 | e.x;
        if (ek == key) return make_uint2(e.z, e.w);
        if (ek == ~0ull) break;
        h = (h + 1) & G.cell_mask;
    }
    return make_uint2(0u, 0u);
}

// Candidates: compaction at min(k + 64, 384) for k <= 64, min(k + 128, 384)
// above (a sweep may ask for k + more, up to 384), plus one chunk of 64.  448 entries keep a 4-wave workgroup's
// LDS at 25.6 KB, so that 6 workgroups fit a CU's 160 KB.
constexpr int kCellCand = 448;

// Workgroups of kCellWaves waves, one query per wave: a workgroup's waves
// are spread over the CU's SIMDs, so a cap of whole workgroups per CU puts
// the same number of waves on every SIMD.  The launch bound asks for 6
// waves per SIMD: the kernel then compiles to 80 VGPRs and 52 bytes of
// scratch per lane (a few values spilled).  At config 4 (r = 0.25) that ran
// 304-306 M queries/s against 285-287 at 5 waves (96 VGPRs, no spill) and
// 295 at 7 (72 VGPRs, 104 bytes spilled); the final gather 228 against 221.
constexpr int kCellWaves = 4;
// Blocks a query may examine before it keeps to whole cells.
constexpr uint32_t kCellMaxBlocks = 256;
// Ranges a query may rank: two per lane (slots lane and lane + 64).  Whole
// cells need at most 64 (4 x 4 x 4); blocks within the radius often number
// 64-100 where the photons are dense, and a query with more keeps to cells.
constexpr uint32_t kCellMaxRanges = 128;
// Up to this many ranges are ranked by counting the smaller keys (a few
// instructions per range); more are sorted (about 27 exchange steps).
constexpr uint32_t kCellRankByCount = 16;

// COUNT: the instrumented copy tallies, per query, the ranges ranked (cells,
// or blocks), the photons read and the candidates (L.tally[0..3]: queries,
// ranges, photons, candidates).
template <bool COUNT>
__global__ __launch_bounds__(64 * kCellWaves, 6) void knn_cell_kernel(KnnLaunch L) {
    __shared__ uint32_t s_key_w[kCellWaves][kCellCand];
    __shared__ uint32_t s_slot_w[kCellWaves][kCellCand];
    __shared__ uint32_t s_hist_w[kCellWaves][256];
    __shared__ uint32_t s_first_w[kCellWaves][kCellMaxRanges], s_pre_w[kCellWaves][kCellMaxRanges];
    __shared__ float s_near_w[kCellWaves][kCellMaxRanges];
    __shared__ uint32_t s_mark_w[kCellWaves][64];
    const uint32_t wv = threadIdx.x / kKnnWave;
    uint32_t* s_key = s_key_w[wv];
    uint32_t* s_slot = s_slot_w[wv];
    uint32_t* s_hist = s_hist_w[wv];
    uint32_t* s_first = s_first_w[wv];
    uint32_t* s_pre = s_pre_w[wv];
    float* s_near = s_near_w[wv];
    uint32_t* s_mark = s_mark_w[wv];
    const PhotonGridView& G = L.grid;
    const uint32_t lane = __lane_id();
    const uint32_t k = L.k;
    // Compaction at k + 64 for k <= 64 (the final sort then holds at most
    // 128 candidates) and k + 128 above (fewer compactions before the final
    // selection: the final gather's k = 100 estimates ran 228 M rays/s at
    // k + 64, 233-235 at k + 128, 232 at k + 192, 227 at k + 256).
    const uint32_t compact_at = min((uint32_t)kCellCand - 64u,
                                    k + (L.compact_slack ? L.compact_slack : (k <= 64u ? 64u : 128u)));
    // Grid-stride over the prepared queries (cell order when sorted): the
    // waves running at one time hold a window of consecutive queries, which
    // share cells.  The next query's record is loaded a query ahead.
    const uint32_t nwaves = gridDim.x * kCellWaves;
    uint32_t qi = blockIdx.x * kCellWaves + wv;
    float4 nq4 = qi < L.nq ? L.qprep[qi] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    uint32_t nq_id = qi < L.nq ? (L.order ? L.order[qi] : qi) : 0u;
    for (; qi < L.nq; qi += nwaves) {
        const uint32_t q = nq_id;
        const float qx = nq4.x, qy = nq4.y, qz = nq4.z;
        const uint32_t cls = __float_as_uint(nq4.w);
        if (qi + nwaves < L.nq) {
            nq4 = L.qprep[qi + nwaves];
            nq_id = L.order ? L.order[qi + nwaves] : qi + nwaves;
        }
        uint32_t cand = 0;
        bool far = false;
        // A query with a coordinate that is not finite has no photon in range
        // (photon positions are finite).
        if (cls != kClassNever && isfinite(qx) && isfinite(qy) && isfinite(qz)) {
            float r2 = L.r2;
            const double qd[3] = {(double)qx, (double)qy, (double)qz};
            // Lane j < 12 works out axis j >> 2 at cell offset j & 3: the
            // cell's coordinate, whether it lies within the grid and the
            // reach, and the squared distance from the query to its widened
            // box.  A lane's own cell is (lane & 3, (lane >> 2) & 3,
            // lane >> 4); it takes its three pairs from those lanes.
            const uint32_t la = min(lane >> 2, 2u);
            const double qa = la == 0 ? qd[0] : la == 1 ? qd[1] : qd[2];
            const double oa = la == 0 ? G.origin[0] : la == 1 ? G.origin[1] : G.origin[2];
            const int32_t dim = (int32_t)(la == 0 ? G.dims[0] : la == 1 ? G.dims[1] : G.dims[2]);
            // widening for assignment and for the rounding of these sums
            const double m = G.cell * 0x1p-20 + (fabs(oa) + fabs(qa)) * 0x1p-48;
            const double w = G.reach + fabs(qa) * 0x1p-40;
            // Cell coordinates (grid_coord) clamped to +-2^30: the grid has
            // fewer than 2^21 cells along an axis, so a clamped coordinate
            // lies outside it either way.
            const int32_t lo = (int32_t)fmin(fmax(floor((qa - w - oa) * G.inv_cell), -0x1p30), 0x1p30);
            const int32_t hi = (int32_t)fmin(fmax(floor((qa + w - oa) * G.inv_cell), -0x1p30), 0x1p30);
            // More than 4 cells only when |q| exceeds 2^38 cells; the grid is
            // refused for maps whose coordinates come near that, so such a
            // query is farther than the reach from every photon.
            far = __ballot(lane < 12u && hi - lo > 3) != 0;
            const int32_t ca = lo + (int32_t)(lane & 3u);
            const uint32_t okm = (uint32_t)__ballot(lane < 12u && ca <= hi && ca >= 0 && ca < dim);
            const double blo = oa + (double)ca * G.cell - m;
            const double bhi = oa + (double)(ca + 1) * G.cell + m;
            const double da = fmax(fmax(blo - qa, qa - bhi), 0.0);
            const double dsq = da * da;
            const uint32_t i0 = lane & 3u, i1 = 4u + ((lane >> 2) & 3u), i2 = 8u + (lane >> 4);
            bool v = ((okm >> i0) & (okm >> i1) & (okm >> i2) & 1u) != 0;
            double near2 = 0.0;
            near2 += __shfl(dsq, (int)i0, 64);
            near2 += __shfl(dsq, (int)i1, 64);
            near2 += __shfl(dsq, (int)i2, 64);
            // the lane's cell (used where it lies within the grid)
            const uint32_t c[3] = {(uint32_t)(__builtin_amdgcn_readlane(lo, 0) + (int32_t)(lane & 3u)),
                                   (uint32_t)(__builtin_amdgcn_readlane(lo, 4) + (int32_t)((lane >> 2) & 3u)),
                                   (uint32_t)(__builtin_amdgcn_readlane(lo, 8) + (int32_t)(lane >> 4))};
            near2 *= 1.0 - 0x1p-20;
            v = v && near2 < (double)r2 && !far;
            uint2 range = make_uint2(0u, 0u), range1 = make_uint2(0u, 0u);  // slots lane, lane + 64
            double near1 = 0.0;
            bool v1 = false;
            bool blk = false;  // the ranges are blocks (wave-uniform)
            uint32_t fb = 0;  // the cell's first block (grids with blocks)
            if (v) {
                if (G.dense) {
                    // below 2^25 entries (the dense table's limit)
                    const uint32_t lin = c[0] + G.dims[0] * (c[1] + G.dims[1] * c[2]);
                    range = G.dense[lin];
                    if (G.blocks) fb = G.dense_block[lin];
                } else {
                    range = cell_lookup(G, cell_key(c[0], c[1], c[2]));
                }
                v = range.y > 0;
            }
            if (G.blocks) {
                // Blocks instead of cells: when the kept cells hold at most
                // kCellMaxBlocks blocks and some cell holds more than one,
                // the wave examines them 64 at a time and keeps those whose
                // box lies within the radius; if at most kCellMaxRanges
                // remain, they replace the cells.  Block j's owner is the last lane whose
                // first block number is at or below j (a lane with no block
                // shares its start with the next lane that has one).
                const uint32_t nb = v ? (range.y + 63u) >> 6 : 0u;
                uint32_t inc = nb;
                for (int o = 1; o < 64; o <<= 1) {
                    const uint32_t t = __shfl_up(inc, o, 64);
                    if (lane >= (uint32_t)o) inc += t;
                }
                const uint32_t nblk = __builtin_amdgcn_readlane(inc, 63);
                if (nblk <= kCellMaxBlocks && nblk > (uint32_t)__popcll(__ballot(v))) {
                    const uint32_t ex = inc - nb;
                    uint32_t nsurv = 0;
                    bool fits = true;
                    for (uint32_t base = 0; base < nblk; base += 64) {
                        const uint32_t j = base + lane;
                        uint32_t o = 0;
                        for (uint32_t step = 32; step; step >>= 1) {
                            const uint32_t e = (uint32_t)__shfl((int)ex, (int)(o + step), 64);
                            if (e <= j) o += step;
                        }
                        const uint32_t b =
                            (uint32_t)__shfl((int)fb, (int)o, 64) + (j - (uint32_t)__shfl((int)ex, (int)o, 64));
                        bool keep = false;
                        float nf = 0.0f;
                        uint32_t bf = 0, bc = 0;
                        if (j < nblk) {
                            const float4 lo = G.blocks[2 * (size_t)b], hi = G.blocks[2 * (size_t)b + 1];
                            // The squared distance to the box with a photon's
                            // float d2 operations: per axis the difference to
                            // the nearer face (the box is the photons' float
                            // min and max, so a photon's difference is at
                            // least as large), squared, summed in the same
                            // order.  Rounding is monotonic, so no photon of
                            // the block has a smaller float d2.
                            const float dx = fmaxf(fmaxf(lo.x - qx, qx - hi.x), 0.0f);
                            const float dy = fmaxf(fmaxf(lo.y - qy, qy - hi.y), 0.0f);
                            const float dz = fmaxf(fmaxf(lo.z - qz, qz - hi.z), 0.0f);
                            const float nr = (dx * dx + dy * dy) + dz * dz;
                            keep = nr < r2;
                            nf = nr;
                            bf = __float_as_uint(lo.w);
                            bc = __float_as_uint(hi.w);
                        }
                        const uint64_t km = __ballot(keep);
                        const uint32_t cnt = (uint32_t)__popcll(km);
                        if (nsurv + cnt > kCellMaxRanges) {
                            fits = false;
                            break;
                        }
                        if (keep) {
                            const uint32_t at = nsurv + lanes_below(km);
                            s_first[at] = bf;
                            s_pre[at] = bc;
                            s_near[at] = nf;
                        }
                        nsurv += cnt;
                    }
                    if (fits) {
                        blk = true;
                        __builtin_amdgcn_wave_barrier();
                        v = lane < nsurv;
                        v1 = lane + 64u < nsurv;
                        if (v) {
                            range = make_uint2(s_first[lane], s_pre[lane]);
                            near2 = (double)s_near[lane];
                        }
                        if (v1) {
                            range1 = make_uint2(s_first[lane + 64u], s_pre[lane + 64u]);
                            near1 = (double)s_near[lane + 64u];
                        }
                        __builtin_amdgcn_wave_barrier();
                    }
                }
            }
            // Nearest first.  The bound is kept as a float rounded down with
            // its 7 low bits cleared (still a lower bound), and the slot in
            // those bits makes the 32-bit sort keys distinct.
            const uint32_t nb = __float_as_uint(__double2float_rd(near2)) & ~127u;
            const uint32_t nb1 = __float_as_uint(__double2float_rd(near1)) & ~127u;
            const uint32_t key = nb | lane, key1 = nb1 | (lane + 64u);
            const uint64_t vm = __ballot(v), vm1 = __ballot(v1);
            const uint32_t ncell = (uint32_t)__popcll(vm) + (uint32_t)__popcll(vm1);
            uint32_t rank = 0, rank1 = 0;
            if (ncell <= kCellRankByCount) {
                // few ranges (slots of the first register only): count the
                // smaller keys
                for (uint64_t rest = vm; rest; rest &= rest - 1) {
                    const uint32_t kj = __builtin_amdgcn_readlane(key, __ffsll((unsigned long long)rest) - 1);
                    rank += kj < key ? 1u : 0u;
                }
            } else {
                // Sort the keys (unused slots as ~0, last); the lane holding
                // sorted position e tells slot (key & 127) its rank e through
                // s_hist, which is free until the first compaction.
                uint32_t x0 = v ? key : 0xFFFFFFFFu, x1 = v1 ? key1 : 0xFFFFFFFFu;
                bitonic_sort(x0, x1, vm1 != 0);
                if (lane < ncell) s_hist[x0 & 127u] = lane;
                if (lane + 64u < ncell) s_hist[x1 & 127u] = lane + 64u;
                __builtin_amdgcn_wave_barrier();
                rank = s_hist[lane];
                rank1 = s_hist[lane + 64u];
                __builtin_amdgcn_wave_barrier();
            }
            if (v) {
                s_first[rank] = range.x;
                s_pre[rank] = range.y;
                s_near[rank] = __uint_as_float(nb);
            }
            if (v1) {
                s_first[rank1] = range1.x;
                s_pre[rank1] = range1.y;
                s_near[rank1] = __uint_as_float(nb1);
            }
            __builtin_amdgcn_wave_barrier();
            uint32_t n_read = 0, n_cand = 0;  // COUNT only
            // Takes the photons of one chunk (position and class, slot) as
            // candidates, and cuts the candidates to the k nearest once k +
            // 64 have gathered (the k-th distance then bounds later ones).
            auto take = [&](const float4& pp, uint32_t ss, bool valid) {
                bool is_c = false;
                float d2 = 0.0f;
                if (valid) {
                    const float x = qx - pp.x, y = qy - pp.y, z = qz - pp.z;
                    d2 = (x * x + y * y) + z * z;
                    is_c = d2 < r2 && __float_as_uint(pp.w) == cls;
                }
                const uint64_t cm = __ballot(is_c);
                if (is_c) {
                    const uint32_t pos = cand + lanes_below(cm);
                    s_key[pos] = __float_as_uint(d2);
                    s_slot[pos] = ss;
                }
                cand += (uint32_t)__popcll(cm);
                if (COUNT) n_cand += (uint32_t)__popcll(cm);
                __builtin_amdgcn_wave_barrier();
                if (cand >= compact_at) {
                    uint32_t need = k;
                    const uint32_t t = select_kth(s_key, cand, need, s_hist);
                    cand = compact_kth(s_key, s_slot, cand, t, need);
                    r2 = __uint_as_float(kept_max(s_key, cand));
                }
            };
            float4 p = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            uint32_t sl = 0;
            if (blk) {
                // Blocks hold at most 64 photons: one block per chunk, lane j
                // taking its photon j; the next block is loaded one ahead.
                auto fetch_block = [&](uint32_t i, float4& fp, uint32_t& fs) -> uint32_t {
                    const uint32_t first = s_first[i], n = s_pre[i];
                    if (lane < n) {
                        fp = G.pos[first + lane];
                        fs = G.slot[first + lane];
                    }
                    return n;
                };
                uint32_t n0 = ncell ? fetch_block(0, p, sl) : 0u;  // no block may lie within the radius
                for (uint32_t i = 0; i < ncell; ++i) {
                    if (s_near[i] >= r2) break;  // this block and all later ones are too far
                    if (COUNT) n_read += n0;
                    float4 pn = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                    uint32_t sn = 0, nn = 0;
                    if (i + 1 < ncell) nn = fetch_block(i + 1, pn, sn);
                    take(p, sl, lane < n0);
                    p = pn;
                    sl = sn;
                    n0 = nn;
                }
            } else {
                // photon counts -> exclusive prefix, in rank order (slots lane,
                // then lane + 64 after the first 64 in total)
                const uint32_t cnt = lane < ncell ? s_pre[lane] : 0u;
                const uint32_t cnt1 = lane + 64u < ncell ? s_pre[lane + 64u] : 0u;
                uint32_t incl = cnt, incl1 = cnt1;
                for (int o = 1; o < 64; o <<= 1) {
                    const uint32_t t = __shfl_up(incl, o, 64), t1 = __shfl_up(incl1, o, 64);
                    if (lane >= (uint32_t)o) {
                        incl += t;
                        incl1 += t1;
                    }
                }
                const uint32_t total0 = __shfl(incl, 63, 64);
                const uint32_t total = total0 + __shfl(incl1, 63, 64);
                __builtin_amdgcn_wave_barrier();
                if (lane < ncell) s_pre[lane] = incl - cnt;
                if (lane + 64u < ncell) s_pre[lane + 64u] = total0 + incl1 - cnt1;
                __builtin_amdgcn_wave_barrier();
                // Chunk e0 covers elements e0 .. e0 + 63 of the concatenated
                // ranges; its photons (position, class, slot) are loaded one
                // chunk ahead, so their latency overlaps the previous chunk.
                //
                // Every kept cell holds a photon, so the starts increase with the
                // cell and an element's cell is the last one starting at or
                // before it.  When few cells start inside a chunk, each lane
                // steps its cell forward; when many do (sparse cells), the chunk
                // marks their starts in LDS and takes a running maximum over the
                // lanes, carried from the previous chunk's last element.
                const uint32_t my_start = lane < ncell ? s_pre[lane] : 0xFFFFFFFFu;
                const uint32_t my_start1 = lane + 64u < ncell ? s_pre[lane + 64u] : 0xFFFFFFFFu;
                uint32_t ci = 0, carry = 0;  // this lane's cell; the cell of the chunk's previous element
                // loads chunk e0; returns the cell of its first element
                auto fetch = [&](uint32_t e0, float4& fp, uint32_t& fs) -> uint32_t {
                    const uint32_t e = e0 + lane;
                    const bool starts_here = lane < ncell && my_start - e0 < 64u;
                    const bool starts_here1 = lane + 64u < ncell && my_start1 - e0 < 64u;
                    if (__popcll(__ballot(starts_here)) + __popcll(__ballot(starts_here1)) <= 3) {
                        if (e < total)
                            while (ci + 1 < ncell && s_pre[ci + 1] <= e) ++ci;
                    } else {
                        s_mark[lane] = 0u;
                        __builtin_amdgcn_wave_barrier();
                        // ranges are not empty: no two start at one element
                        if (starts_here) s_mark[my_start - e0] = lane;
                        if (starts_here1) s_mark[my_start1 - e0] = lane + 64u;
                        __builtin_amdgcn_wave_barrier();
                        uint32_t c = max(s_mark[lane], carry);
                        for (int o = 1; o < 64; o <<= 1) {
                            const uint32_t t = __shfl_up(c, o, 64);
                            if (lane >= (uint32_t)o) c = max(c, t);
                        }
                        ci = c;
                    }
                    carry = __shfl(ci, 63, 64);
                    if (e < total) {
                        const uint32_t gi = s_first[ci] + (e - s_pre[ci]);
                        fp = G.pos[gi];
                        fs = G.slot[gi];
                    }
                    return __shfl(ci, 0, 64);
                };
                uint32_t c0 = total ? fetch(0, p, sl) : 0u;  // the cell of the chunk's first element
                for (uint32_t e0 = 0; e0 < total; e0 += 64) {
                    if (s_near[c0] >= r2) break;  // this cell and all later ones are too far
                    if (COUNT) n_read += min(64u, total - e0);
                    float4 pn = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                    uint32_t sn = 0;
                    uint32_t c0n = c0;
                    if (e0 + 64 < total) c0n = fetch(e0 + 64, pn, sn);
                    take(p, sl, e0 + lane < total);
                    p = pn;
                    sl = sn;
                    c0 = c0n;
                }
            }
            if (COUNT && lane == 0) {
                atomicAdd(&L.tally[1], (unsigned long long)ncell);
                atomicAdd(&L.tally[2], (unsigned long long)n_read);
                atomicAdd(&L.tally[3], (unsigned long long)n_cand);
            }
        }
        if (COUNT && lane == 0) atomicAdd(&L.tally[0], 1ull);
        wave_finish(L, q, cand, false, s_key, s_slot, s_hist);
        __builtin_amdgcn_wave_barrier();
    }
}

// Prepared queries for knn_cell_kernel: position and normal class in the
// order it serves them.
__global__ void knn_cell_prep(KnnLaunch L, float4* prep) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= L.nq) return;
    const uint32_t q = L.order ? L.order[i] : i;
    const uint32_t cls = lookup_class(L.table, L.table_mask, L.qnrm[3 * (size_t)q], L.qnrm[3 * (size_t)q + 1],
                                      L.qnrm[3 * (size_t)q + 2]);
    prep[i] = make_float4(L.qpos[3 * (size_t)q], L.qpos[3 * (size_t)q + 1], L.qpos[3 * (size_t)q + 2],
                          __uint_as_float(cls));
}

// Queries of the tree walks in the order of a 30-bit Morton code of their
// position within the map's box (clamped to it; ~0 for a position that is
// not finite), so that the lanes of a wave walk nearby parts of the tree.
__global__ void knn_tree_keys(KnnLaunch L, float3 lo, float3 inv, uint32_t* keys, uint32_t* ids) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= L.nq) return;
    const float x[3] = {L.qpos[3 * (size_t)i], L.qpos[3 * (size_t)i + 1], L.qpos[3 * (size_t)i + 2]};
    uint32_t key = 0xFFFFFFFFu;
    if (isfinite(x[0]) && isfinite(x[1]) && isfinite(x[2])) {
        const float o[3] = {lo.x, lo.y, lo.z}, s[3] = {inv.x, inv.y, inv.z};
        uint32_t c[3];
        for (int a = 0; a < 3; ++a) c[a] = (uint32_t)fminf(fmaxf((x[a] - o[a]) * s[a], 0.0f), 1023.0f);
        key = (uint32_t)grid_spread21(c[0]) | (uint32_t)grid_spread21(c[1]) << 1 | (uint32_t)grid_spread21(c[2]) << 2;
    }
    keys[i] = key;
    ids[i] = i;
}

__global__ void knn_cell_keys(KnnLaunch L, unsigned long long* keys, uint32_t* ids) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= L.nq) return;
    const PhotonGridView& G = L.grid;
    unsigned long long key = ~0ull;
    const float x[3] = {L.qpos[3 * (size_t)i], L.qpos[3 * (size_t)i + 1], L.qpos[3 * (size_t)i + 2]};
    if (isfinite(x[0]) && isfinite(x[1]) && isfinite(x[2])) {
        uint64_t c[3];
        for (int a = 0; a < 3; ++a) {
            const int64_t v = grid_coord((double)x[a], G.origin[a], G.inv_cell);
            c[a] = (uint64_t)min(max(v, (int64_t)0), (int64_t)G.dims[a] - 1);
        }
        key = cell_key(c[0], c[1], c[2]);
    }
    keys[i] = key;
    ids[i] = i;
}

}  // namespace dev

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

size_t knn_cell_order_bytes(uint32_t nq) {
    size_t tmp = 0;
    RQ_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, (unsigned long long*)nullptr,
                                              (unsigned long long*)nullptr, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                              (int)nq, 0, 64));
    return 2 * align256((size_t)nq * 8) + align256((size_t)nq * 4) + align256(std::max<size_t>(tmp, 16));
}

void knn_cell_order(const KnnLaunch& L, uint32_t* order, void* scratch, size_t scratch_bytes, hipStream_t s) {
    const uint32_t nq = L.nq;
    char* w = static_cast<char*>(scratch);
    auto* keys = reinterpret_cast<unsigned long long*>(w);
    auto* skeys = reinterpret_cast<unsigned long long*>(w + align256((size_t)nq * 8));
    auto* ids = reinterpret_cast<uint32_t*>(w + 2 * align256((size_t)nq * 8));
    void* tmp = w + 2 * align256((size_t)nq * 8) + align256((size_t)nq * 4);
    size_t tmp_bytes = scratch_bytes - (size_t)(static_cast<char*>(tmp) - w);
    hipLaunchKernelGGL(dev::knn_cell_keys, dim3((nq + 255) / 256), dim3(256), 0, s, L, keys, ids);
    RQ_HIP(hipGetLastError());
    // Cell coordinates are clamped into the grid (fewer than 2^b cells per
    // axis), so Morton keys use the low 3b bits; a query that is not finite
    // has the key ~0, whose low 3b bits are all set, and sorts last either
    // way.  Sorting those bits only takes 3b / 8 passes rather than 8.
    uint32_t b = 1;
    const uint32_t widest = std::max(std::max(L.grid.dims[0], L.grid.dims[1]), L.grid.dims[2]);
    while (b < 21 && (1u << b) < widest) ++b;
    RQ_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, keys, skeys, ids, order, (int)nq, 0, (int)(3 * b), s));
}

size_t knn_tree_order_bytes(uint32_t nq) {
    size_t tmp = 0;
    RQ_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                              (uint32_t*)nullptr, (uint32_t*)nullptr, (int)nq, 0, 32));
    return 3 * align256((size_t)nq * 4) + align256(std::max<size_t>(tmp, 16));
}

void knn_tree_order(const KnnLaunch& L, const float bmin[3], const float bmax[3], uint32_t* order, void* scratch,
                    size_t scratch_bytes, hipStream_t s) {
    const uint32_t nq = L.nq;
    char* w = static_cast<char*>(scratch);
    auto* keys = reinterpret_cast<uint32_t*>(w);
    auto* skeys = reinterpret_cast<uint32_t*>(w + align256((size_t)nq * 4));
    auto* ids = reinterpret_cast<uint32_t*>(w + 2 * align256((size_t)nq * 4));
    void* tmp = w + 3 * align256((size_t)nq * 4);
    size_t tmp_bytes = scratch_bytes - (size_t)(static_cast<char*>(tmp) - w);
    float inv[3];
    for (int a = 0; a < 3; ++a) {
        const float ext = bmax[a] - bmin[a];
        inv[a] = ext > 0.0f && std::isfinite(ext) ? 1024.0f / ext : 0.0f;
    }
    hipLaunchKernelGGL(dev::knn_tree_keys, dim3((nq + 255) / 256), dim3(256), 0, s, L,
                       make_float3(bmin[0], bmin[1], bmin[2]), make_float3(inv[0], inv[1], inv[2]), keys, ids);
    RQ_HIP(hipGetLastError());
    RQ_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, keys, skeys, ids, order, (int)nq, 0, 30, s));
}

uint32_t knn_wave_stack_cap(uint32_t n) {
    uint32_t d = 1;  // levels of the implicit balanced tree over n photons
    while ((n >> d) != 0) ++d;
    return 128u * (d + 2);
}

int knn_wave_grid(uint32_t nq, int num_cus) {
    int per_cu = 0;
    RQ_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)dev::knn_wave_kernel, dev::kKnnWave, 0));
    return (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)nq, (int64_t)std::max(1, per_cu) * num_cus));
}

void launch_knn_wave(const KnnLaunch& L, int grid, uint32_t stack_cap, hipStream_t s) {
    if (L.nq == 0) return;
    hipLaunchKernelGGL(dev::knn_wave_kernel, dim3(grid), dim3(dev::kKnnWave), 0, s, L, stack_cap);
    RQ_HIP(hipGetLastError());
}

int knn_cell_grid(uint32_t nq, int num_cus) {
    int per_cu = 0;
    RQ_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)dev::knn_cell_kernel<false>,
                                                        dev::kKnnWave * dev::kCellWaves, 0));
    // Six waves per SIMD (24 per CU, as the launch bound compiles for).
    // With one-wave workgroups, before the kernel's blocks and its present
    // register count, the 4M-query final gather ran 160 M rays/s at 5 waves
    // per SIMD, 146-149 M when 20 workgroups per CU could land 6, 6, 6, 2 on
    // the SIMDs, and 113 M with 24 per CU.  RQ_KNN_CELL_WAVES sets the count
    // for sweeps.
    int cap = 24;
    if (const char* e = std::getenv("RQ_KNN_CELL_WAVES")) cap = std::atoi(e);
    per_cu = std::max(1, std::min(per_cu, std::max(1, cap / dev::kCellWaves)));
    const int64_t need = ((int64_t)nq + dev::kCellWaves - 1) / dev::kCellWaves;
    return (int)std::max<int64_t>(1, std::min<int64_t>(need, (int64_t)per_cu * num_cus));
}

void launch_knn_cells(const KnnLaunch& L, float4* prep, int grid, hipStream_t s) {
    if (L.nq == 0) return;
    hipLaunchKernelGGL(dev::knn_cell_prep, dim3((L.nq + 255) / 256), dim3(256), 0, s, L, prep);
    RQ_HIP(hipGetLastError());
    KnnLaunch C = L;
    C.qprep = prep;
    // RQ_KNN_COMPACT_SLACK: compact the candidates once they exceed k by this
    // many (default 64, or 128 for k > 64), for sweeps; results do not depend
    // on it.
    if (const char* e = std::getenv("RQ_KNN_COMPACT_SLACK")) C.compact_slack = (uint32_t)std::max(1, std::atoi(e));
    if (C.tally)
        hipLaunchKernelGGL(dev::knn_cell_kernel<true>, dim3(grid), dim3(dev::kKnnWave * dev::kCellWaves), 0, s, C);
    else
        hipLaunchKernelGGL(dev::knn_cell_kernel<false>, dim3(grid), dim3(dev::kKnnWave * dev::kCellWaves), 0, s, C);
    RQ_HIP(hipGetLastError());
}

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
