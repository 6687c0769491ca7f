// Velocity sums of the interior faces formed per cell (gfx950): the K1
// sums of the interior-face pipeline (k_flux_split.h, calc_flux_mtforcev/h,
// 2d/src/flux.f90:68-102,116-120,184-202) from one read of each cell's
// upwind arrays for both face orientations, instead of one read per
// orientation.
//
// A face's sums are sums over its nodes of (weight x polynomial in vn, vt)
// x an upwind value, taken from the left/lower cell where
// vn - fn dt/2 >= 0 (upwind_delta) and from the right/upper cell
// elsewhere.  Split instead by the sign of vn alone, a cell's nodes fall
// into two half-spaces per axis: u >= 0 feeds its right face (the cell is
// its left cell), u < 0 its left face, and likewise along v.  The sums are
// linear in the upwind values, so a face's sums are the sum of its left
// cell's u >= 0 (or v >= 0) bin and its right cell's u < 0 (v < 0) bin
// whenever no node changes side between vn and vn - fn dt/2 ("flips").
//
//   k_cell_sum    one wavefront per cell, in place of k_slope2d: the van
//                 Leer slopes as k_slope2d forms and stores them, then the
//                 four half-space bins of the cell's 37 sums, in FP64,
//                 from h, b and the slopes in LDS
//   k_fface_cs    one lane per interior face: the face's sums as the sum
//                 of two bins, or, where a node flips, K1's own pass over
//                 the face (fsum_face, by the lane's wavefront), then the
//                 face state, Maxwellian sums and conserved flux as k_fface
//
// Lane mapping of k_cell_sum: lane l takes column iu = l % nv of row
// jv = j0 + l / nv (64 / nv rows per pass; nv <= 64), so each lane keeps
// one u for the whole cell and the rows of a pass share their v.  With the
// weights a tensor product, w = wx(u) wy(v), the vertical-face sums factor
// as wx u^a x (sums over rows of wy v^b x value) and the horizontal-face
// ones as wx (-u)^b x (sums of wy v^a x value): 16 and 18 FP64
// accumulators per lane instead of 2 x 35 sums, expanded into the 37 sums
// once per bin and reduced over the wavefront.  Rows are taken in two
// passes, v < 0 then v >= 0 (rows [0, js) and [js, nv), checked on the
// host to split the v nodes that way), so a pass's rows all feed one
// horizontal bin.  The rows go through LDS a chunk at a time: 16-byte
// loads of the five stencil cells' h and b, the slopes formed and stored in
// that layout, then h, b and the slopes written to LDS and read back in the
// column layout.  The 16-byte loads keep several KB per wavefront in flight
// where the column layout's own loads would keep a few hundred bytes.  At
// 512^2 cells, GH40 (K1 of the split pipeline: 4.25 ms FP32, 6.19 FP64;
// the slope kernel 2.47 / 5.01):
//   sums reading the stored slopes, one node per load   3.6 ms FP32
//   the same staged through LDS                          2.5 / 3.9 ms
//   slopes formed here, no slope kernel                  4.0 / 5.9 ms
#pragma once
#include "hugks_device.h"
#include "hugks_kernels.h"
#include "k_flux2d.h"
#include "k_flux_split.h"

namespace hugks {

// bins of a cell: 0 u < 0, 1 u >= 0, 2 v < 0, 3 v >= 0 (as upwind_delta),
// stored by cell: csum[(cell * CSUM_BINS + bin) * CELL_SUM_LD + q] (stored
// by sum instead, with the cells fastest, the bin stores of the cell
// kernel scatter: slopes and sums 3.71 -> 3.86 ms FP32, 5.9 -> 6.4 ms FP64
// at 512^2 cells, GH40)
constexpr int CSUM_BINS = 4;
__host__ __device__ inline size_t csum_at(int bin, int q, size_t cell) {
    return (cell * CSUM_BINS + (size_t)bin) * CELL_SUM_LD + (size_t)q;
}

// The 37 sums of a vertical face (vn = u, vt = v) from a lane's
// accumulators over its column: x[0..3] = sum a v^b, x[4..6] = e v^b,
// x[7..10] = f v^b, x[11..12] = g v^b, x[13] = p, x[14..15] = r v^b, with
// a = w hf, e = w sn, f = w st, g = w bf, p = w bn, r = w bt as in
// fsum_face.  Zero when !on.
__device__ __forceinline__ void cs_expand_v(const double (&x)[16], double u, bool on,
                                            double (&s)[FSUM_N]) {
    const double A0 = x[0], A1 = x[1], A2 = x[2], A3 = x[3];
    const double E0 = x[4], E1 = x[5], E2 = x[6];
    const double F0 = x[7], F1 = x[8], F2 = x[9], F3 = x[10];
    const double G0 = x[11], G1 = x[12], P0 = x[13], R0 = x[14], R1 = x[15];
    const double u2 = u * u, u3 = u2 * u;
    const double e2 = u2 * E0 + E2, f1 = u2 * F1 + F3, f0 = u2 * F0 + F2;
    s[0] = A0;
    s[1] = u * A0;
    s[2] = A1;
    s[4] = G0;
    s[5] = F0;
    s[6] = u * F0;
    s[7] = F1;
    s[8] = f0;
    s[9] = R0;
    s[10] = u2 * E0;
    s[11] = u * F1;
    s[12] = u2 * A0;
    s[13] = u3 * E0;
    s[14] = u2 * F1;
    s[15] = u * A1;
    s[16] = u2 * E1;
    s[17] = u * F2;
    s[19] = u * G0;
    s[20] = u2 * e2;
    s[21] = u2 * P0;
    s[22] = u * f1;
    s[23] = u * R1;
    s[24] = u * E0;
    s[25] = u * E1;
    s[26] = u * e2;
    s[27] = u * P0;
    s[28] = u2 * F0;
    s[29] = u * f0;
    s[30] = u * R0;
    s[31] = A2;
    s[32] = u3 * A0;
    s[33] = u2 * A1;
    s[34] = u * A2;
    s[35] = A3;
    s[36] = G1;
    s[3] = s[12] + s[31];
    s[18] = s[32] + s[34];
#pragma unroll
    for (int q = 0; q < FSUM_N; ++q) s[q] = on ? s[q] : 0.0;
}

// The 37 sums of a horizontal face (vn = v, vt = -u), tangential-slope
// sums negated as fsum_face does, from x[0..3] = sum a v^b, x[4..7] =
// e v^(b+1), x[8..11] = f v^b, x[12..13] = g v^b, x[14..15] = p v^(b+1),
// x[16..17] = r v^b.
__device__ __forceinline__ void cs_expand_h(const double (&x)[18], double u,
                                            double (&s)[FSUM_N]) {
    const double A0 = x[0], A1 = x[1], A2 = x[2], A3 = x[3];
    const double E1 = x[4], E2 = x[5], E3 = x[6], E4 = x[7];
    const double F0 = x[8], F1 = x[9], F2 = x[10], F3 = x[11];
    const double G0 = x[12], G1 = x[13], P1 = x[14], P2 = x[15], R0 = x[16], R1 = x[17];
    const double u2 = u * u, nu = -u;
    const double f1 = F3 + u2 * F1;
    s[0] = A0;
    s[1] = A1;
    s[2] = nu * A0;
    s[4] = G0;
    s[5] = F0;
    s[6] = F1;
    s[7] = nu * F0;
    s[8] = F2 + u2 * F0;
    s[9] = R0;
    s[10] = E2;
    s[11] = nu * F1;
    s[12] = A2;
    s[13] = E3;
    s[14] = nu * F2;
    s[15] = nu * A1;
    s[16] = nu * E2;
    s[17] = u2 * F1;
    s[19] = G1;
    s[20] = E4 + u2 * E2;
    s[21] = P2;
    s[22] = nu * f1;
    s[23] = nu * R1;
    s[24] = E1;
    s[25] = nu * E1;
    s[26] = E3 + u2 * E1;
    s[27] = P1;
    s[28] = F2;
    s[29] = f1;
    s[30] = R1;
    s[31] = u2 * A0;
    s[32] = A3;
    s[33] = nu * A2;
    s[34] = u2 * A1;
    s[35] = nu * u2 * A0;
    s[36] = nu * G0;
    s[3] = s[12] + s[31];
    s[18] = s[32] + s[34];
    // horizontal faces take the tangential slopes with the opposite sign
    constexpr int tq[13] = {5, 6, 7, 8, 9, 11, 14, 17, 22, 23, 28, 29, 30};
#pragma unroll
    for (int t = 0; t < 13; ++t) s[tq[t]] = -s[tq[t]];
}

// One bin: the 37 sums expanded from a lane's accumulators (vertical V,
// else horizontal) and reduced over the wavefront in two halves (fewer
// values live at once), written to out[q * stride], q = 0..36.
template <bool V, int N>
__device__ __forceinline__ void cs_reduce_bin(const double (&x0)[N], double u, double wu, bool on,
                                              double* __restrict__ out, size_t stride) {
    // the column's weight factor, left out of the per-node products
    double x[N];
#pragma unroll
    for (int i = 0; i < N; ++i) x[i] = wu * x0[i];
    constexpr int H1 = 19, H2 = FSUM_N - H1;
#pragma unroll
    for (int half = 0; half < 2; ++half) {
        double s[FSUM_N];
        if constexpr (V) {
            static_assert(N == 16, "vertical accumulators");
            cs_expand_v(x, u, on, s);
        } else {
            static_assert(N == 18, "horizontal accumulators");
            cs_expand_h(x, u, s);
        }
        int q;
        if (half == 0) {
            double t[H1];
#pragma unroll
            for (int i = 0; i < H1; ++i) t[i] = s[i];
            const double y = wave_reduce_scatter<H1>(t, q);
            if (q >= 0) out[(size_t)q * stride] = y;
        } else {
            double t[H2];
#pragma unroll
            for (int i = 0; i < H2; ++i) t[i] = s[H1 + i];
            const double y = wave_reduce_scatter<H2>(t, q);
            if (q >= 0) out[(size_t)(H1 + q) * stride] = y;
        }
    }
}

// The van Leer slope of vl_slope.  FP32 takes the limiter on the raw
// differences, (sgn dR + sgn dL) |dR| |dL| / (|dR| + |dL| + smv len) / len,
// the same quantity with the 1/len factor applied once instead of to both
// differences (one VALU fewer per slope; the rounding differs), and the
// quotient from the hardware reciprocal.  FP64 is vl_slope's own arithmetic.
template <typename A>
__device__ __forceinline__ A cs_slope(A fl, A fn, A fr, bool inner, A inv_den, A smv, A smv_len) {
    if constexpr (sizeof(A) == 4) {
        if (inner) {
            const A dL = fn - fl, dR = fr - fn;
            const A num = (copysign((A)1, dR) + copysign((A)1, dL)) * fabs(dR) * fabs(dL);
            const A den = fabs(dR) + fabs(dL) + smv_len;
            return fast_div(num, den) * inv_den;
        }
        return (fr - fl) * inv_den;
    } else {
        return vl_slope<A, true>(fl, fn, fr, inner, inv_den, smv);
    }
}

// Bytes of one LDS chunk per array and wavefront: whole rows, at most
// what one 16-byte load per lane covers.
constexpr int CS_CHUNK_BYTES = WAVE * 16;
constexpr int CS_MAX_ROWS = 16;
__host__ __device__ constexpr int cs_chunk_rows(int nv, int es) {
    return CS_CHUNK_BYTES / (nv * es) < 1 ? 1
         : CS_CHUNK_BYTES / (nv * es) > CS_MAX_ROWS ? CS_MAX_ROWS
         : CS_CHUNK_BYTES / (nv * es);
}
// LDS of one wavefront: the chunk of the six arrays, then each chunk row's
// weight factor times v^k, k = 0..4 (FP64, six to a row for 16-byte reads)
constexpr int CS_ROWC = 6;
template <typename T>
constexpr int cs_lds_per_wave() {
    return 6 * CS_CHUNK_BYTES + CS_ROWC * CS_MAX_ROWS * 8;
}

// The slopes and the four half-space bins of one cell.  The rows are
// staged through LDS a chunk at a time: every lane first loads its 16-byte
// pieces of the chunk from the five stencil cells (VW elements; one element
// per load where a row is not a multiple of 16 bytes), forms and stores the
// slopes and puts h, b and the slopes in LDS, then the lanes read the chunk
// back one node per lane in the column layout above.
template <typename T, int VW>
__device__ __forceinline__ void k_cell_sum_body(
    Consts c, const DevState* __restrict__ st, const T* __restrict__ h, const T* __restrict__ b,
    T* __restrict__ shx, T* __restrict__ shy, T* __restrict__ sbx, T* __restrict__ sby,
    const double* __restrict__ u1, const double* __restrict__ v1,
    const double* __restrict__ wx, const double* __restrict__ wy, double* __restrict__ csum,
    int js, unsigned char* __restrict__ lds, int bid, int nblk) {
    using A = typename Arith<T>::A;
    if (st->stopped) return;
    const int lane = threadIdx.x & (WAVE - 1);
    const int L = __builtin_amdgcn_readfirstlane(
        (int)xcd_block(bid, nblk) * SLOPE_CPW + (int)(threadIdx.x / WAVE));
    if (L >= c.nx * c.ny) return;
    const int cell = tiled_cell(L, c.nx, c.ny);
    const size_t o = (size_t)cell * c.ldv;
    // the slope stencils, as k_slope2d takes them
    const A smv = (A)c.smv;
    bool inner[2];
    A inv_den[2], smv_len[2];
    size_t on[2][2];
#pragma unroll
    for (int dir = 0; dir < 2; ++dir) {
        const SlopeNb nb = slope_nb(c, cell, dir);
        const A len = (A)(dir == 0 ? c.dx : c.dy);
        inner[dir] = nb.inner;
        inv_den[dir] = (A)1 / ((A)0.5 * len + (A)0.5 * len);
        smv_len[dir] = smv / inv_den[dir];
        on[dir][0] = (size_t)nb.cl * c.ldv;
        on[dir][1] = (size_t)nb.cr * c.ldv;
    }
    const int nv = c.nv;
    const int CR = cs_chunk_rows(nv, (int)sizeof(T));
    // LDS elements per array: the fixed chunk size, so that the row loop
    // reads the six arrays from one address with immediate offsets
    constexpr int cw = CS_CHUNK_BYTES / (int)sizeof(T);
    unsigned char* myb = lds + (size_t)(threadIdx.x / WAVE) * cs_lds_per_wave<T>();
    T* my = reinterpret_cast<T*>(myb);
    double* rowc = reinterpret_cast<double*>(myb + 6 * CS_CHUNK_BYTES);
    const int R = WAVE / nv;                 // rows per pass of the sums
    const int r = lane / nv, iu = lane - r * nv;
    const bool act = r < R;
    const double u = act ? u1[iu] : 0.0;
    const double wu = act ? wx[iu] : 0.0;    // the column's weight factor
    const int su = upwind_delta((A)u);
    const double hlx = 0.5 * c.dx, hly = 0.5 * c.dy;
    const double sgnV = su ? hlx : -hlx;
    // av: the vertical faces' accumulators over the whole column; per pass
    // (v < 0, then v >= 0) the column's row moments w v^k of h, b and the
    // slopes, from which the pass's horizontal accumulators and its share of
    // av follow (the face values h +- len/2 slope are linear in them):
    // qh[k] = sum h c_k (k < 4), mx = sx (k < 4), my = sy (k < 5),
    // mb = b (k < 2), mtx = tx (k < 2), mty = ty (k < 3)
    double av[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) av[q] = 0.0;
    typedef T Vec __attribute__((ext_vector_type(VW)));
#pragma unroll 1
    for (int p = 0; p < 2; ++p) {
        const int lo = p ? js : 0, hi = p ? nv : js;
        const double sgnH = p ? hly : -hly;
        double qh[4], qx[4], qy[5], qb[2], qtx[2], qty[3];
#pragma unroll
        for (int k = 0; k < 4; ++k) qh[k] = qx[k] = 0.0;
#pragma unroll
        for (int k = 0; k < 5; ++k) qy[k] = 0.0;
#pragma unroll
        for (int k = 0; k < 2; ++k) qb[k] = qtx[k] = 0.0;
#pragma unroll
        for (int k = 0; k < 3; ++k) qty[k] = 0.0;
#pragma unroll 1
        for (int j0 = lo; j0 < hi; j0 += CR) {
            const int cr = hi - j0 < CR ? hi - j0 : CR;
            const int ne = cr * nv;          // elements of this chunk per array
            const int base = j0 * nv;
            // every load of the chunk first, then the LDS stores
            constexpr int NL = VW > 1 ? 1 : (CS_CHUNK_BYTES / (int)sizeof(T) + WAVE - 1) / WAVE;
            double vr = 0.0, wr = 0.0;
#pragma unroll
            for (int t = 0; t < NL; ++t) {
                const int e0 = (lane + WAVE * t) * VW;
                const bool in = e0 < ne;
                const size_t g = o + (size_t)base + (size_t)(in ? e0 : 0);
                // lanes past the chunk load its first piece again (in
                // bounds) and do not store it
                const Vec x0 = *reinterpret_cast<const Vec*>(h + g);
                const Vec x1 = *reinterpret_cast<const Vec*>(b + g);
                // the slopes from the five-cell stencils, as k_slope2d forms
                // them but with the limiter's quotient from the hardware
                // reciprocal for FP32 (within 1 ulp; the IEEE division was
                // a third of this kernel's VALU instructions: 4.0 -> 3.8 ms
                // at 512^2 cells, GH40), stored as k_slope2d stores them
                Vec x2, x3, x4, x5;
                {
                    const size_t gr = g - o;
                    const Vec hxl = *reinterpret_cast<const Vec*>(h + on[0][0] + gr);
                    const Vec hxr = *reinterpret_cast<const Vec*>(h + on[0][1] + gr);
                    const Vec hyl = *reinterpret_cast<const Vec*>(h + on[1][0] + gr);
                    const Vec hyr = *reinterpret_cast<const Vec*>(h + on[1][1] + gr);
                    const Vec bxl = *reinterpret_cast<const Vec*>(b + on[0][0] + gr);
                    const Vec bxr = *reinterpret_cast<const Vec*>(b + on[0][1] + gr);
                    const Vec byl = *reinterpret_cast<const Vec*>(b + on[1][0] + gr);
                    const Vec byr = *reinterpret_cast<const Vec*>(b + on[1][1] + gr);
#pragma unroll
                    for (int q = 0; q < VW; ++q) {
                        x2[q] = (T)cs_slope<A>((A)hxl[q], (A)x0[q], (A)hxr[q], inner[0],
                                               inv_den[0], smv, smv_len[0]);
                        x3[q] = (T)cs_slope<A>((A)hyl[q], (A)x0[q], (A)hyr[q], inner[1],
                                               inv_den[1], smv, smv_len[1]);
                        x4[q] = (T)cs_slope<A>((A)bxl[q], (A)x1[q], (A)bxr[q], inner[0],
                                               inv_den[0], smv, smv_len[0]);
                        x5[q] = (T)cs_slope<A>((A)byl[q], (A)x1[q], (A)byr[q], inner[1],
                                               inv_den[1], smv, smv_len[1]);
                    }
                    if (in) {
                        // streamed out: read back only by the face kernels
                        __builtin_nontemporal_store(x2, reinterpret_cast<Vec*>(shx + g));
                        __builtin_nontemporal_store(x3, reinterpret_cast<Vec*>(shy + g));
                        __builtin_nontemporal_store(x4, reinterpret_cast<Vec*>(sbx + g));
                        __builtin_nontemporal_store(x5, reinterpret_cast<Vec*>(sby + g));
                    }
                }
                if (t == 0) {
                    // the chunk rows' v and weight factor, with the loads
                    const int jr = lane < cr ? j0 + lane : j0;
                    vr = v1[jr];
                    wr = wy[jr];
                    wave_lds_sync();     // the previous chunk's reads are done
                }
                if (in) {
                    *reinterpret_cast<Vec*>(my + e0) = x0;
                    *reinterpret_cast<Vec*>(my + cw + e0) = x1;
                    *reinterpret_cast<Vec*>(my + 2 * cw + e0) = x2;
                    *reinterpret_cast<Vec*>(my + 3 * cw + e0) = x3;
                    *reinterpret_cast<Vec*>(my + 4 * cw + e0) = x4;
                    *reinterpret_cast<Vec*>(my + 5 * cw + e0) = x5;
                }
            }
            // the rows' weight factor times powers of their v, formed once
            // per row here rather than by every lane of the row loop
            if (lane < cr) {
                double* rc = rowc + CS_ROWC * lane;
                const double c1 = wr * vr, c2 = c1 * vr, c3 = c2 * vr;
                rc[0] = wr;
                rc[1] = c1;
                rc[2] = c2;
                rc[3] = c3;
                rc[4] = c3 * vr;
            }
            wave_lds_sync();
            const int jn = act ? cr : 0;
#pragma unroll 1
            for (int jj = r; jj < jn; jj += R) {
                const int e = jj * nv + iu;
                // the row's weight factor times powers of its v
                const double* rc = rowc + CS_ROWC * jj;
                const double c0 = rc[0], c1 = rc[1], c2 = rc[2], c3 = rc[3], c4 = rc[4];
                const double hu = (double)my[e], bu = (double)my[cw + e];
                const double sx = (double)my[2 * cw + e], sy = (double)my[3 * cw + e];
                const double tx = (double)my[4 * cw + e], ty = (double)my[5 * cw + e];
                qh[0] += hu * c0;
                qh[1] += hu * c1;
                qh[2] += hu * c2;
                qh[3] += hu * c3;
                qx[0] += sx * c0;
                qx[1] += sx * c1;
                qx[2] += sx * c2;
                qx[3] += sx * c3;
                qy[0] += sy * c0;
                qy[1] += sy * c1;
                qy[2] += sy * c2;
                qy[3] += sy * c3;
                qy[4] += sy * c4;
                qb[0] += bu * c0;
                qb[1] += bu * c1;
                qtx[0] += tx * c0;
                qtx[1] += tx * c1;
                qty[0] += ty * c0;
                qty[1] += ty * c1;
                qty[2] += ty * c2;
            }
        }
        // the horizontal bin of this pass's rows, from the horizontal
        // faces' accumulators (a = h + sgnH sy, g = b + sgnH ty)
        {
            const double ah[18] = {qh[0] + sgnH * qy[0], qh[1] + sgnH * qy[1],
                                   qh[2] + sgnH * qy[2], qh[3] + sgnH * qy[3],
                                   qy[1], qy[2], qy[3], qy[4],
                                   qx[0], qx[1], qx[2], qx[3],
                                   qb[0] + sgnH * qty[0], qb[1] + sgnH * qty[1],
                                   qty[1], qty[2], qtx[0], qtx[1]};
            cs_reduce_bin<false>(ah, u, wu, true, csum + csum_at(2 + p, 0, cell), 1);
        }
        // this pass's share of the vertical accumulators (a = h + sgnV sx,
        // g = b + sgnV tx)
#pragma unroll
        for (int k = 0; k < 4; ++k) av[k] += qh[k] + sgnV * qx[k];
#pragma unroll
        for (int k = 0; k < 3; ++k) av[4 + k] += qx[k];
#pragma unroll
        for (int k = 0; k < 4; ++k) av[7 + k] += qy[k];
        av[11] += qb[0] + sgnV * qtx[0];
        av[12] += qb[1] + sgnV * qtx[1];
        av[13] += qtx[0];
        av[14] += qty[0];
        av[15] += qty[1];
    }
#pragma unroll 1
    for (int side = 0; side < 2; ++side)
        cs_reduce_bin<true>(av, u, wu, act && su == side, csum + csum_at(side, 0, cell), 1);
}
#ifndef HUGKS_BODIES_ONLY
template <typename T, int VW>
__global__ __launch_bounds__(WAVE * SLOPE_CPW) HUGKS_WAVES_PER_EU(3) void k_cell_sum(
    Consts c, const DevState* __restrict__ st, const T* __restrict__ h, const T* __restrict__ b,
    T* __restrict__ shx, T* __restrict__ shy, T* __restrict__ sbx, T* __restrict__ sby,
    const double* __restrict__ u1, const double* __restrict__ v1,
    const double* __restrict__ wx, const double* __restrict__ wy, double* __restrict__ csum,
    int js) {
    __shared__ __align__(16) unsigned char lds[SLOPE_CPW * cs_lds_per_wave<T>()];
    k_cell_sum_body<T, VW>(c, st, h, b, shx, shy, sbx, sby, u1, v1, wx, wy, csum, js, lds,
                           (int)blockIdx.x, (int)gridDim.x);
}
#endif

// Whether a node of face fg changes upwind side between vn and
// vn - fn dt/2 (a "flip"): then the face's sums are not the sum of two
// bins.  The test runs in the arithmetic type, as the upwind choice of
// fsum_face and fnode_face does.
template <typename A>
__device__ __forceinline__ bool face_flips(const Consts& c, const FaceGeo& fg,
                                           const double* __restrict__ u1,
                                           const double* __restrict__ v1,
                                           const double* __restrict__ acc_v,
                                           const double* __restrict__ acc_h, double dt) {
    const double* acc = fg.dir == 0 ? acc_v : acc_h;
    const double* x = fg.dir == 0 ? u1 : v1;
    const A vs = (A)(0.5 * acc[2 * fg.face] * dt);
    bool f = false;
    for (int i = 0; i < c.nv; ++i) {
        const A xi = (A)x[i];
        f = f || upwind_delta(xi - vs) != upwind_delta(xi);
    }
    return f;
}

// K2, K3 (tensor-product quadrature) and K4 of one face on one lane, as
// k_fface, with the face's 37 sums taken first: its left cell's forward
// bin plus its right cell's backward bin, or, for a face with a flip,
// K1's pass over the face (fsum_face) by the whole wavefront, for each
// such face of the wavefront in turn.
template <typename T>
__device__ __forceinline__ void k_fface_cs_body(
    Consts c, const DevState* __restrict__ st, const double* __restrict__ W,
    const T* __restrict__ h, const T* __restrict__ b, const T* __restrict__ shx,
    const T* __restrict__ shy, const T* __restrict__ sbx, const T* __restrict__ sby,
    const double* __restrict__ u1, const double* __restrict__ v1,
    const double* __restrict__ wx, const double* __restrict__ wy,
    const double* __restrict__ und, const double* __restrict__ vnd,
    const double* __restrict__ wnd, const double* __restrict__ node4,
    const double* __restrict__ acc_v, const double* __restrict__ acc_h,
    const double* __restrict__ csum, double* __restrict__ fsum, double* __restrict__ fcoef,
    double* __restrict__ vflux, double* __restrict__ hflux, int nvi, int nint, int bid) {
    using A = typename Arith<T>::A;
    if (st->stopped) return;
    const int g = bid * 256 + (int)threadIdx.x;
    const bool have = g < nint;
    const double dt = st->dt;
    FaceGeo fg{};
    bool flip = false;
    if (have) {
        fg = face_geo(c, g, nvi);
        flip = face_flips<A>(c, fg, u1, v1, acc_v, acc_h, dt);
    }
    // the wavefront's faces with a flip, one after the other, by all lanes
    unsigned long long fm = __ballot(flip);
    if (fm != 0) {
        while (fm != 0) {
            const int l = __builtin_ctzll(fm);
            fm &= fm - 1;
            const int gf = __shfl(g, l, WAVE);
            const FaceGeo ff = face_geo(c, gf, nvi);
            const double* acc = ff.dir == 0 ? acc_v : acc_h;
            const double vshift = 0.5 * acc[2 * ff.face] * dt;
            const Upwind<T, false> up = make_upwind<T, false>(c, ff, h, b, shx, shy, sbx, sby);
            const NodeTabs nt{und, vnd, wnd, node4};
            if (ff.dir == 0)
                fsum_face<0, T, false>(c, vshift, up, nt, fsum + (size_t)gf * FSUM_LD);
            else
                fsum_face<1, T, false>(c, vshift, up, nt, fsum + (size_t)gf * FSUM_LD);
        }
        __threadfence();   // those sums, written by other lanes, before their faces read them
    }
    if (!have) return;
    if (!flip) {
        const double* bL = csum + csum_at(fg.dir ? 3 : 1, 0, (size_t)fg.cL);
        const double* bR = csum + csum_at(fg.dir ? 2 : 0, 0, (size_t)fg.cR);
        double* out = fsum + (size_t)g * FSUM_LD;
#pragma unroll
        for (int q = 0; q < FSUM_N; ++q) out[q] = bL[q] + bR[q];
    }
    fstate_lane(c, dt, W, acc_v, acc_h, fsum, fcoef, fg, g);
    double s2[FMAXW_N];
    fmaxw_sep_lane(c, u1, v1, wx, wy, acc_v, acc_h, fcoef, fg, g, s2);
    fslope_t_lane(c, acc_v, acc_h, s2, fcoef, vflux, hflux, fg, g);
}
#ifndef HUGKS_BODIES_ONLY
template <typename T>
__global__ __launch_bounds__(256) void k_fface_cs(
    Consts c, const DevState* __restrict__ st, const double* __restrict__ W,
    const T* __restrict__ h, const T* __restrict__ b, const T* __restrict__ shx,
    const T* __restrict__ shy, const T* __restrict__ sbx, const T* __restrict__ sby,
    const double* __restrict__ u1, const double* __restrict__ v1,
    const double* __restrict__ wx, const double* __restrict__ wy,
    const double* __restrict__ und, const double* __restrict__ vnd,
    const double* __restrict__ wnd, const double* __restrict__ node4,
    const double* __restrict__ acc_v, const double* __restrict__ acc_h,
    const double* __restrict__ csum, double* __restrict__ fsum, double* __restrict__ fcoef,
    double* __restrict__ vflux, double* __restrict__ hflux, int nvi, int nint) {
    k_fface_cs_body<T>(c, st, W, h, b, shx, shy, sbx, sby, u1, v1, wx, wy, und, vnd, wnd, node4,
                       acc_v, acc_h, csum, fsum, fcoef, vflux, hflux, nvi, nint,
                       (int)blockIdx.x);
}
#endif

}  // namespace hugks
