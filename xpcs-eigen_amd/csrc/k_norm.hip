// (d) Partition reduction and g2 normalisation (Corr::normalizeG2s,
// src/xpcs/corr.cpp:927-1091) on the exact per-pixel sums.
//
//   k_norm_partial  one workgroup per chunk (<= 256 pixels of one (q, s) group):
//                   integer sums of G2/IP/IF per tau (flat-field weighted sums in
//                   FP64 when a flat field is set) and a Welford of
//                   x_p = N G2_p / (IP_p IF_p) with 0/0 -> 0 over each of the
//                   chunk's two pixel slices, the two merged in order.
//   k_norm_group    per (group, tau): sums of the group's chunks in order;
//                   g2s = c_s N sum G2 / (sum IP sum IF)            (FP64)
//   k_norm_q        per (q, tau): mean of the non-NaN g2s of q; the chunk
//                   Welford states merged in order (Chan et al.), giving
//                   stderr = sqrt(1/n) sqrt(M2/n) with every pixel of q counted.
// The reduction order is fixed, so results are deterministic.
#include "device_util.h"
#include "internal.h"

namespace xpcs {

constexpr int kNormBatch = 8;
constexpr int kNormBlock = 256;
constexpr int kNormSlices = 2;                        // pixel slices of a chunk (2: measured best of 1, 2, 4, 8)
constexpr int kNormLanes = kNormBlock / kNormSlices;  // taus per pass

// Running state of one (chunk slice, tau): the sums and a Welford of x_p.
struct NormAcc {
    int64_t sg = 0, sp = 0, sf = 0;
    double wg = 0.0, wp = 0.0, wf = 0.0;
    double n = 0.0, mean = 0.0, m2 = 0.0;
};

// How a normalisation reads its sums: integers without weights (the exact
// path), integers with flat-field weights (FP64 weighted sums), or FP64 sums
// (frame-sum normalised values or reference floats; weights optional).  The
// unweighted integer path carries no FP64 sums (94 VGPRs, five waves per
// SIMD; it reads at the same rate as with three: the kernel is bound by the
// HBM reads of the sums).
enum NormMode { kNormInt = 0, kNormWeighted = 1, kNormFp = 2 };

// Pixels [pb, pe) of tau column i in order. kNormFp: the sums are FP64 stored
// in the same 8-byte buffers; they take the weighted (FP64) route with weight
// 1 when the plan has no flat field.  W32: integer IP and IF in 32-bit words,
// with high words at hw + idx for the pixels whose total passes 2^32 - 1.
// The spread of x_p over the slice is kept as sums of x - K and (x - K)^2
// about the slice's first value K (no division per pixel besides x_p's own;
// the slice is short, so the shifted sums lose nothing a Welford update
// would keep) and leaves as (n, mean, M2) for the ordered merges.
template <int MODE, bool W32>
__device__ __forceinline__ void norm_accumulate(const int64_t* __restrict__ G2, const int64_t* __restrict__ IP,
                                                const int64_t* __restrict__ IF, const double* __restrict__ weight,
                                                const uint64_t* __restrict__ pix_total, int64_t hw, int32_t S,
                                                int i, double N, int pb, int pe, NormAcc& a)
{
    double K = 0.0, s1 = 0.0, s2 = 0.0;
    // kNormBatch pixels' loads in flight per thread
    for (int p0 = pb; p0 < pe; p0 += kNormBatch) {
        int64_t gg[kNormBatch], aa[kNormBatch], bb[kNormBatch];
        uint64_t tmax = 0;
#pragma unroll
        for (int u = 0; u < kNormBatch; u++) {
            const int pu = min(p0 + u, pe - 1);
            const int64_t idx = (int64_t)pu * S + i;
            gg[u] = G2[idx];
            if constexpr (W32) {
                aa[u] = (int64_t)reinterpret_cast<const uint32_t*>(IP)[idx];
                bb[u] = (int64_t)reinterpret_cast<const uint32_t*>(IF)[idx];
                tmax = max(tmax, pix_total[pu]);
            } else {
                aa[u] = IP[idx];
                bb[u] = IF[idx];
            }
        }
        if constexpr (W32) {
            // pixels with high words (totals of 2^32 or more) are rare
            if (tmax > 0xFFFFFFFFull) {
#pragma unroll
                for (int u = 0; u < kNormBatch; u++) {
                    const int pu = min(p0 + u, pe - 1);
                    const int64_t idx = (int64_t)pu * S + i;
                    const int64_t hi = sum_hi(pix_total[pu], hw);
                    aa[u] = get_sum(IP, idx, true, hi);
                    bb[u] = get_sum(IF, idx, true, hi);
                }
            }
        }
        const int nu = min(kNormBatch, pe - p0);
#pragma unroll
        for (int u = 0; u < kNormBatch; u++) {
            if (u >= nu) continue;
            const int p = p0 + u;
            const int64_t g = gg[u];
            const int64_t ip = aa[u];
            const int64_t f = bb[u];
            double x;
            if constexpr (MODE == kNormFp) {
                const double w = weight ? weight[p] : 1.0;
                const double gd = __longlong_as_double(g), ad = __longlong_as_double(ip),
                             bd = __longlong_as_double(f);
                a.wg += w * w * gd;
                a.wp += w * ad;
                a.wf += w * bd;
                x = (w == 0.0) ? 0.0 : N * gd / (ad * bd);
            } else if constexpr (MODE == kNormWeighted) {
                const double w = weight[p];
                a.wg += w * w * (double)g;
                a.wp += w * (double)ip;
                a.wf += w * (double)f;
                x = (w == 0.0) ? 0.0 : N * (double)g / ((double)ip * (double)f);
            } else {
                a.sg += g;
                a.sp += ip;
                a.sf += f;
                x = N * (double)g / ((double)ip * (double)f);
            }
            if (x != x) x = 0.0;
            if (p == pb) K = x;
            const double d = x - K;
            s1 += d;
            s2 += d * d;
        }
    }
    const double n = (double)max(pe - pb, 0);
    a.n = n;
    if (n > 0.0) {
        a.mean = K + s1 / n;
        a.m2 = fmax(s2 - s1 * (s1 / n), 0.0);
    }
}

// a followed by b: sums added, Welford states merged (Chan et al.).
template <int MODE>
__device__ __forceinline__ void norm_merge(NormAcc& a, const NormAcc& b)
{
    if constexpr (MODE == kNormInt) {
        a.sg += b.sg;
        a.sp += b.sp;
        a.sf += b.sf;
    } else {
        a.wg += b.wg;
        a.wp += b.wp;
        a.wf += b.wf;
    }
    if (b.n == 0.0) return;
    if (a.n == 0.0) {
        a.n = b.n;
        a.mean = b.mean;
        a.m2 = b.m2;
        return;
    }
    const double n = a.n + b.n;
    const double d = b.mean - a.mean;
    a.mean += d * (b.n / n);
    a.m2 += b.m2 + d * d * (a.n * b.n / n);
    a.n = n;
}

// Each chunk's pixels are split into kNormSlices consecutive slices, one
// thread per (slice, tau) for each pass of up to kNormLanes taus, and the
// slice states are merged in slice order through LDS. A workgroup takes
// kNormBlock / (kNormSlices * lanes) chunks at once (several when the tau
// range is narrow, as in stream snapshots). Neither the tau range nor the
// number of chunks per workgroup changes a chunk's computation, so a range of
// taus gets the same partials as the full schedule.
template <int MODE, bool W32>
__global__ void __launch_bounds__(kNormBlock)
k_norm_partial(const NormChunk* __restrict__ chunks, int32_t nchunks, const TauParam* __restrict__ taus,
               int32_t t0, int32_t t1, int32_t S, const int64_t* __restrict__ G2, const int64_t* __restrict__ IP,
               const int64_t* __restrict__ IF, const double* __restrict__ weight,
               const uint64_t* __restrict__ pix_total, int64_t hw, NormPartial* __restrict__ part)
{
    __shared__ __align__(16) unsigned char s_raw[kNormBlock * sizeof(NormAcc)];
    NormAcc* s_acc = reinterpret_cast<NormAcc*>(s_raw);
    const int nt = t1 - t0;
    const int lanes = min(nt, kNormLanes);              // taus per pass
    const int per_chunk = kNormSlices * lanes;
    const int cpb = max(1, kNormBlock / per_chunk);     // chunks per workgroup (the launch sizes the block to it)
    const int cc = (int)threadIdx.x / per_chunk;
    const int r = (int)threadIdx.x % per_chunk;
    const int sl = r / lanes, lt = r % lanes;
    for (int64_t c0 = (int64_t)blockIdx.x * cpb; c0 < nchunks; c0 += (int64_t)gridDim.x * cpb) {
        const int64_t c = c0 + cc;
        const bool live = cc < cpb && c < nchunks;
        NormChunk ch{0, 0};
        if (live) ch = chunks[c];
        const int len = ch.end - ch.begin;
        const int pb = ch.begin + (int)((int64_t)len * sl / kNormSlices);
        const int pe = ch.begin + (int)((int64_t)len * (sl + 1) / kNormSlices);
        for (int tb = t0; tb < t1; tb += lanes) {   // uniform trip count
            const int i = tb + lt;
            const bool act = live && i < t1;
            if (act) {
                NormAcc a;
                norm_accumulate<MODE, W32>(G2, IP, IF, weight, pix_total, hw, S, i, (double)taus[i].N, pb, pe, a);
                s_acc[threadIdx.x] = a;
            }
            __syncthreads();
            if (act && sl == 0) {
                NormAcc a = s_acc[threadIdx.x];
#pragma unroll
                for (int k = 1; k < kNormSlices; k++) norm_merge<MODE>(a, s_acc[threadIdx.x + k * lanes]);
                NormPartial o;
                constexpr bool fpsum = MODE != kNormInt;
                o.g2 = fpsum ? __double_as_longlong(a.wg) : a.sg;
                o.ip = fpsum ? __double_as_longlong(a.wp) : a.sp;
                o.ifs = fpsum ? __double_as_longlong(a.wf) : a.sf;
                o.n = a.n;
                o.mean = a.mean;
                o.m2 = a.m2;
                part[c * S + i] = o;
            }
            __syncthreads();
        }
    }
}

// One wavefront per (group, tau): lane l sums the group's chunk partials
// l, l + 64, ...; the lanes then reduce in a fixed tree (integer sums are exact
// in any order; weighted FP64 sums come out the same on every run).
__global__ void __launch_bounds__(64)
k_norm_group(const NormPartial* __restrict__ part, const int32_t* __restrict__ group_chunk_first,
             const int32_t* __restrict__ group_size, int32_t ngroups, const TauParam* __restrict__ taus,
             int32_t t0, int32_t t1, int32_t S, int weighted, double* __restrict__ g2s)
{
    const int64_t k = blockIdx.x;
    const int nt = t1 - t0;
    if (k >= (int64_t)ngroups * nt) return;
    const int g = (int)(k / nt);
    const int i = t0 + (int)(k % nt);
    const int lane = (int)threadIdx.x;
    const int c0 = group_chunk_first[g], c1 = group_chunk_first[g + 1];
    double sg, sp, sf;
    if (weighted) {
        double a = 0.0, b = 0.0, c = 0.0;
        for (int ch = c0 + lane; ch < c1; ch += 64) {
            const NormPartial& r = part[(int64_t)ch * S + i];
            a += __longlong_as_double(r.g2);
            b += __longlong_as_double(r.ip);
            c += __longlong_as_double(r.ifs);
        }
        for (int o = 32; o > 0; o >>= 1) {
            a += __shfl_xor(a, o, 64);
            b += __shfl_xor(b, o, 64);
            c += __shfl_xor(c, o, 64);
        }
        sg = a;
        sp = b;
        sf = c;
    } else {
        int64_t a = 0, b = 0, c = 0;
        for (int ch = c0 + lane; ch < c1; ch += 64) {
            const NormPartial& r = part[(int64_t)ch * S + i];
            a += r.g2;
            b += r.ip;
            c += r.ifs;
        }
        for (int o = 32; o > 0; o >>= 1) {
            a += __shfl_xor(a, o, 64);
            b += __shfl_xor(b, o, 64);
            c += __shfl_xor(c, o, 64);
        }
        sg = (double)a;
        sp = (double)b;
        sf = (double)c;
    }
    if (lane == 0) g2s[(int64_t)g * S + i] = (double)group_size[g] * (double)taus[i].N * sg / (sp * sf);
}

// Chan's merge of Welford states (n, mean, m2); an empty state is the identity.
__device__ __forceinline__ void welford_merge(double& n, double& mean, double& m2, double rn, double rmean,
                                              double rm2)
{
    if (rn == 0.0) return;
    if (n == 0.0) {
        n = rn;
        mean = rmean;
        m2 = rm2;
        return;
    }
    const double nt = n + rn;
    const double d = rmean - mean;
    mean += d * (rn / nt);
    m2 += rm2 + d * d * (n * rn / nt);
    n = nt;
}

// One wavefront per (q, tau): g2 = mean of the non-NaN static-bin values;
// stderr from the chunk Welford states, merged lane-strided (lane l takes
// chunks l, l + 64, ... in order) and then across lanes in a fixed tree.
__global__ void __launch_bounds__(64)
k_norm_q(const NormPartial* __restrict__ part, const double* __restrict__ g2s,
         const int32_t* __restrict__ q_first_group, const int32_t* __restrict__ q_chunk_first,
         int32_t nq, int32_t t0, int32_t t1, int32_t S, double* __restrict__ g2, double* __restrict__ se)
{
    const int64_t k = blockIdx.x;
    const int nt = t1 - t0;
    if (k >= (int64_t)nq * nt) return;
    const int q = (int)(k / nt);    // 0-based
    const int i = t0 + (int)(k % nt);
    const int lane = (int)threadIdx.x;
    if (lane == 0) {
        double acc = 0.0, cnt = 0.0;
        for (int g = q_first_group[q]; g < q_first_group[q + 1]; g++) {
            const double v = g2s[(int64_t)g * S + i];
            if (v == v) {
                acc += v;
                cnt += 1.0;
            }
        }
        g2[(int64_t)i * nq + q] = acc / cnt;
    }
    double n = 0.0, mean = 0.0, m2 = 0.0;
    for (int c = q_chunk_first[q] + lane; c < q_chunk_first[q + 1]; c += 64) {
        const NormPartial& r = part[(int64_t)c * S + i];
        welford_merge(n, mean, m2, r.n, r.mean, r.m2);
    }
    for (int o = 1; o < 64; o <<= 1) {
        const double rn = __shfl_xor(n, o, 64), rmean = __shfl_xor(mean, o, 64), rm2 = __shfl_xor(m2, o, 64);
        // the lower lane of each pair keeps the order (lower chunks first)
        if ((lane & o) == 0) welford_merge(n, mean, m2, rn, rmean, rm2);
        else {
            double n2 = rn, mean2 = rmean, m22 = rm2;
            welford_merge(n2, mean2, m22, n, mean, m2);
            n = n2;
            mean = mean2;
            m2 = m22;
        }
    }
    if (lane == 0) se[(int64_t)i * nq + q] = sqrt(1.0 / n) * sqrt(m2 / n);
}

// Per-pixel sums in the reference's float units, layout [T][W*H] (FP: the
// sums are FP64 bit patterns).
__device__ __forceinline__ double sum_value(int64_t v, bool fp) { return fp ? __longlong_as_double(v) : (double)v; }

template <bool FP>
__global__ void __launch_bounds__(256)
k_ref_f32(const int64_t* __restrict__ G2, const int64_t* __restrict__ IP, const int64_t* __restrict__ IF,
          bool w32, const uint64_t* __restrict__ pix_total,
          const int32_t* __restrict__ c2pix, int32_t P, const TauParam* __restrict__ taus, int32_t T,
          const double* __restrict__ weight, double inv_scale, int64_t npix_det,
          float* __restrict__ oG2, float* __restrict__ oIP, float* __restrict__ oIF)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (int64_t)P * T) return;
    const int i = (int)(k / P);
    const int64_t c = k % P;
    const TauParam tp = taus[i];
    const double s = (weight ? weight[c] : 1.0) * inv_scale;
    const double lv = (double)(1ll << tp.level);
    const double N = (double)tp.N;
    const int64_t o = (int64_t)i * npix_det + c2pix[c];
    oG2[o] = (float)(sum_value(G2[c * T + i], FP) * s * s / (lv * lv * N));
    const int64_t hi = w32 ? sum_hi(pix_total[c], (int64_t)P * T) : 0;
    oIP[o] = (float)(sum_value(get_sum(IP, c * T + i, w32, hi), FP) * s / (lv * N));
    oIF[o] = (float)(sum_value(get_sum(IF, c * T + i, w32, hi), FP) * s / (lv * N));
}

// Reference-unit float arrays [T][W*H] (the G2s/IPs/IFs of a caller, as
// Corr::normalizeG2s receives them) as FP64 sums in summed-bin units, [P][T]
// in compact pixel order: G2 = v 4^L N, IP = v 2^L N, IF = v 2^L N.
__global__ void __launch_bounds__(256)
k_from_ref_f32(const float* __restrict__ G2, const float* __restrict__ IP, const float* __restrict__ IF,
               const int32_t* __restrict__ c2pix, int32_t P, const TauParam* __restrict__ taus, int32_t T,
               int64_t npix_det, int64_t* __restrict__ oG2, int64_t* __restrict__ oIP, int64_t* __restrict__ oIF)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (int64_t)P * T) return;
    const int64_t c = k / T;
    const int i = (int)(k % T);
    const TauParam tp = taus[i];
    const double lv = (double)(1ll << tp.level);
    const double N = (double)tp.N;
    const int64_t src = (int64_t)i * npix_det + c2pix[c];
    oG2[k] = __double_as_longlong((double)G2[src] * (lv * lv * N));
    oIP[k] = __double_as_longlong((double)IP[src] * (lv * N));
    oIF[k] = __double_as_longlong((double)IF[src] * (lv * N));
}

// Normalisation of taus [t0, t1) of the sums G2/IP/IF: rows of stride S
// ([P][S]), int64 integers (w32: IP and IF in the 32-bit layout, the plan's
// pixel totals telling which rows have high words) or (fp) FP64 bit patterns, each pixel weighted by
// weight[c] (NULL: 1); taus gives N per tau; the chunk partials and group
// values use the same stride, g2 and stderr are [tau][q].  Values of other
// taus are left as they are.  The partials, group values and results go to
// `out` (the plan's buffers, or caller / scratch memory).
void launch_normalize_sums(xpcs_mt_plan* p, hipStream_t s, int32_t t0, int32_t t1, int32_t S,
                           const TauParam* taus, const int64_t* G2, const int64_t* IP, const int64_t* IF, bool fp,
                           bool w32, const double* weight, const NormOut& out)
{
    const HostTables& h = p->h;
    const int32_t nchunks = (int32_t)h.chunks.size();
    const int32_t nt = t1 - t0;
    if (nt <= 0) return;
    timer_begin(p, s, "k_norm");
    // chunks per workgroup as the kernel derives them from the tau range
    const int cpb = std::max(1, kNormBlock / (kNormSlices * std::min(nt, kNormLanes)));
    const int64_t nwg = (nchunks + cpb - 1) / cpb;
    const unsigned nb = nwg < 65536 ? (unsigned)nwg : 65536u;
    // only the threads the chunks use, whole waves (T = 88: 192 threads, not 256)
    const int used = cpb * kNormSlices * std::min(nt, kNormLanes);
    const int nthr = std::min(kNormBlock, (used + 63) / 64 * 64);
    if (nchunks > 0) {
        auto k = fp       ? k_norm_partial<kNormFp, false>
                 : weight ? (w32 ? k_norm_partial<kNormWeighted, true> : k_norm_partial<kNormWeighted, false>)
                 : (w32 ? k_norm_partial<kNormInt, true> : k_norm_partial<kNormInt, false>);
        k<<<nb, nthr, 0, s>>>(p->d_chunks, nchunks, taus, t0, t1, S, G2, IP, IF, weight, p->d_pix_total,
                              (int64_t)h.P * S, out.part);
    }
    const int64_t ng = (int64_t)h.ngroups * nt;
    if (ng > 0)
        k_norm_group<<<(unsigned)ng, 64, 0, s>>>(out.part, p->d_group_chunk_first, p->d_group_size, h.ngroups,
                                                 taus, t0, t1, S, (fp || weight) ? 1 : 0, out.g2s);
    const int64_t nqt = (int64_t)h.nq * nt;
    if (nqt > 0)
        k_norm_q<<<(unsigned)nqt, 64, 0, s>>>(out.part, out.g2s, p->d_q_first_group, p->d_q_chunk_first,
                                              h.nq, t0, t1, S, out.g2, out.se);
    timer_end(p, s);
}

NormOut plan_norm_out(xpcs_mt_plan* p) { return NormOut{p->d_part, p->d_g2s, p->d_g2, p->d_se}; }

void launch_normalize_with(xpcs_mt_plan* p, hipStream_t s, int32_t t0, int32_t t1, int32_t S,
                           const TauParam* taus)
{
    launch_normalize_sums(p, s, t0, t1, S, taus, p->d_G2, p->d_IP, p->d_IF, p->fsnorm, p->ipif32 && !p->fsnorm,
                          p->d_weight, plan_norm_out(p));
}

void launch_from_ref_f32(xpcs_mt_plan* p, hipStream_t s, const float* G2, const float* IP, const float* IF,
                         int64_t* oG2, int64_t* oIP, int64_t* oIF)
{
    const HostTables& h = p->h;
    const int64_t n = (int64_t)h.P * h.T;
    if (n == 0) return;
    k_from_ref_f32<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(G2, IP, IF, p->d_c2pix, h.P, p->d_taus, h.T,
                                                               (int64_t)h.W * h.H, oG2, oIP, oIF);
}

void launch_normalize(xpcs_mt_plan* p, hipStream_t s)
{
    launch_normalize_with(p, s, 0, p->h.T, p->h.T, p->d_taus);
}

void launch_ref_f32(xpcs_mt_plan* p, hipStream_t s, float* G2, float* IP, float* IF)
{
    const HostTables& h = p->h;
    const int64_t n = (int64_t)h.P * h.T;
    const double inv_scale = 1.0 / ((double)h.avg * (double)p->nscale);
    auto k = p->fsnorm ? k_ref_f32<true> : k_ref_f32<false>;
    k<<<(int)((n + 255) / 256), 256, 0, s>>>(p->d_G2, p->d_IP, p->d_IF, p->ipif32 && !p->fsnorm, p->d_pix_total,
                                             p->d_c2pix, h.P, p->d_taus, h.T,
                                             p->d_weight, inv_scale, (int64_t)h.W * h.H, G2, IP, IF);
}

}  // namespace xpcs
