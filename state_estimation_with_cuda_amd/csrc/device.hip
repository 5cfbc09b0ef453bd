// device.hip -- gfx950 kernels and the device-resident Gauss-Newton loop.
//
// Precision (the mixed form): h(x), residuals, Huber weights, the right-hand
// side H^T W r, the state and the convergence test are FP64; the gain is
// formed and factored in FP64 on the iterations that refactor, and the
// factor's solve data (panels, subtree solve matrices, top inverse blocks) are
// rounded once to FP32 for the iterations on the cached factor (T = float in
// the solve kernels), which then read half the bytes.  The Gauss-Newton loop
// corrects the rounding like any other inexact step.  The plain form
// instantiates the same kernels with T = double.  Option: an FP32
// factorization instead (mixed64 off), whose pivots below threshold switch the
// instance to the FP64 factor.
//
// Kernel map (reference counterparts, replaced rather than translated):
//   k_ybus        fillYbusKernel / initYbusDiagonalKernel, sparse_matrix.cu:389-492
//   k_power       computePowerInjectionsKernel, kernels.cu:129-230
//   k_flows       computeBranchFlowsKernel, kernels.cu:241-333
//   k_meas        computeMeasurementFunctionKernel + computeResidualsKernel +
//                 computeObjectiveFunctionKernel + computeHuberWeightsKernel +
//                 computeJacobianValuesKernel + applyWeightScalingKernel
//                 (kernels.cu:381-549, 634-671; sparse_matrix.cu:905-926,
//                 1627-2147), fused into one pass over the measurement rows
//   k_spgemm      cusparseSpGEMM G = H^T W H, sparse_matrix.cu:1209-1356
//   k_factor_*    cusolverSpXcsrlsvchol factor, sparse_matrix.cu:856-882
//   k_fwd / k_bwd cusolverSpXcsrlsvchol solve (same call)
//   k_rhs         computeRHSVectorKernel, kernels.cu:823-862
//   k_update      applyStateUpdateKernel + computeMaxMismatchKernel +
//                 checkValidityKernel, kernels.cu:721-782, 926-941
//   k_flat        flatStartKernel, kernels.cu:799-817
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <map>
#include <mutex>
#include <vector>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>

#include "instance.h"

namespace sle {

namespace {

struct HipError : GpuError {
    explicit HipError(const std::string& s) : GpuError(s) {}
};

#define HIPCHECK(x)                                                                                     \
    do {                                                                                                \
        hipError_t e_ = (x);                                                                            \
        if (e_ != hipSuccess)                                                                           \
            throw HipError(std::string("HIP error ") + hipGetErrorString(e_) + " at " + #x + " (" + \
                           __FILE__ + ":" + std::to_string(__LINE__) + ")");                            \
    } while (0)

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kFrontA = 32;    // fronts with c <= 32: LDS front of 32x32
constexpr int kFrontB = 64;    // 32 < c <= 64: LDS front of 64x64
constexpr int kVecMax = 2048;  // triangular-solve LDS vector bound for large fronts
constexpr int kSpinLimit = 1 << 22;  // polls before a dependency wait gives up (~0.1 s)
constexpr double kNewtonSwitch = 1e-2;  // Huber: step size below which the outliers' gain weights drop
constexpr double kRefreshRatio = 0.25;  // REALTIME: step ratio above which the cached factor is refreshed
constexpr double kLinearRegime = 1e-3;  // steps below this: Gauss-Newton contracts fast with a sound factor
// an iteration on the cached factor is queued behind the running one unless
// the running one's predicted step is below kAheadMargin * tol
constexpr double kAheadMargin = 2.0;

// measurement code byte: type (4 bits) | end << 4 | active << 5 | from-first << 6
// (from-first: the branch's from bus precedes its to bus in the natural state
// order, which fixes the column order of a flow row)
__host__ __device__ inline uint8_t meas_code(int type, int end, int active, int from_first) {
    return (uint8_t)((type & 15) | ((end & 1) << 4) | ((active & 1) << 5) | ((from_first & 1) << 6));
}

// Workgroups of a launch are dealt to the 8 XCDs round-robin and every XCD
// has its own L2.  A kernel may take its logical block from xcd_block(): XCD x
// then runs the contiguous logical range [x*q + min(x, r), +q + (x < r)) of
// the nb = 8q + r blocks, so blocks that touch neighbouring buses share one L2.
// Measured on the 200k-bus grid it pays for the power and right-hand-side sums
// only: the measurement pass and the products of the solves run faster in
// plain block order (their scattered writes and the 256 MB last-level cache
// shared by all XCDs favour neighbouring blocks running at the same time).
constexpr int kXcds = 8;
__constant__ int c_xcd_remap = 1;

__device__ inline int xcd_block() {
    const int b = blockIdx.x;
    if (!c_xcd_remap) return b;
    const int nb = gridDim.x, x = b % kXcds, l = b / kXcds, q = nb / kXcds, r = nb % kXcds;
    return x * q + (x < r ? x : r) + l;
}

struct DevStatus {
    double maxdx;
    double J;
    double rmax;
    int ridx;
    int nan;
    int piv;
    int flow_err;
    uint32_t seq;  // iterations finished (k_maxfinish); the host copy's is written last
    // set by k_maxfinish when an iteration converged (or produced a NaN step)
    // while a next iteration may already be queued behind it: that iteration
    // then leaves the state and the status alone.  Cleared by the pass that
    // opens a solve's first iteration (k_power, or k_meas_loc with reset).
    int stop;
};

// ---------------------------------------------------------------------------
// wave64 / block reductions
// ---------------------------------------------------------------------------
__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}
__device__ inline double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, kWave));
    return v;
}
// (value, index) max with ties to the lower index
__device__ inline void wave_argmax(double& v, int& i) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        double ov = __shfl_xor(v, o, kWave);
        int oi = __shfl_xor(i, o, kWave);
        if (ov > v || (ov == v && oi >= 0 && (i < 0 || oi < i))) {
            v = ov;
            i = oi;
        }
    }
}

// ---------------------------------------------------------------------------
// Ybus values from the per-entry branch contribution lists (FP64, no atomics)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_ybus(int nY, const int32_t* __restrict__ ycptr,
                                                 const int32_t* __restrict__ ycon, const double* __restrict__ gs,
                                                 const double* __restrict__ bs, const double* __restrict__ bc,
                                                 const double* __restrict__ tap, const double* __restrict__ shift,
                                                 const uint8_t* __restrict__ status, double* __restrict__ yG,
                                                 double* __restrict__ yB) {
    int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= nY) return;
    double gr = 0.0, gi = 0.0;
    for (int p = ycptr[e]; p < ycptr[e + 1]; ++p) {
        int con = ycon[p];
        int k = con >> 2, kind = con & 3;
        if (!status[k]) continue;
        double g = gs[k], b = bs[k], a = tap[k], hb = 0.5 * bc[k];
        double a2 = a * a;
        if (kind == 0) {
            gr += g / a2;
            gi += (b + hb) / a2;
        } else if (kind == 1) {
            gr += g;
            gi += b + hb;
        } else {
            double cp = cos(shift[k]), sp = sin(shift[k]);
            if (kind == 2) {
                gr += -(g * cp - b * sp) / a;
                gi += -(g * sp + b * cp) / a;
            } else {
                gr += -(g * cp + b * sp) / a;
                gi += -(b * cp - g * sp) / a;
            }
        }
    }
    yG[e] = gr;
    yB[e] = gi;
}

// Ybus values at the transposed positions: entry e of row j (column l) gets
// Y_lj, read by k_rhs_loc in its row walk; refreshed after every k_ybus.
__global__ __launch_bounds__(kBlock) void k_ybus_tr(int nY, const int32_t* __restrict__ ytr,
                                                   const double* __restrict__ yG, const double* __restrict__ yB,
                                                   double* __restrict__ yGt, double* __restrict__ yBt) {
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= nY) return;
    const int f = ytr[e];
    yGt[e] = yG[f];
    yBt[e] = yB[f];
}

// ---------------------------------------------------------------------------
// Bus injections; per Ybus entry P_ij, Q_ij terms kept for the Jacobian
// ---------------------------------------------------------------------------
constexpr int kPowerChunk = 4;  // Ybus entries of a row loaded together

__global__ __launch_bounds__(kBlock) void k_power(int n, const int32_t* __restrict__ yrow,
                                                  const int32_t* __restrict__ ycol, const double* __restrict__ yG,
                                                  const double* __restrict__ yB, const double* __restrict__ V,
                                                  const double* __restrict__ th, double* __restrict__ P,
                                                  double* __restrict__ Q, double* __restrict__ Pij,
                                                  double* __restrict__ Qij, DevStatus* st) {
    int i = xcd_block() * kBlock + threadIdx.x;
    if (i == 0) {
        st->nan = 0;
        st->piv = 0;
        st->flow_err = 0;
        st->stop = 0;
    }
    if (i >= n) return;
    double vi = V[i], ti = th[i];
    double p = 0.0, q = 0.0;
    // the row in chunks of kPowerChunk entries, loads issued together (indices
    // clamped to the row's last entry); one sincos in the loop body, the chunk
    // shifting down a register per entry
    const int e1 = yrow[i + 1];
    for (int e = yrow[i]; e < e1; e += kPowerChunk) {
        int jj[kPowerChunk];
        double G[kPowerChunk], Bq[kPowerChunk], Vq[kPowerChunk], tq[kPowerChunk];
#pragma unroll
        for (int u = 0; u < kPowerChunk; ++u) {
            const int ee = min(e + u, e1 - 1);
            jj[u] = ycol[ee];
            G[u] = yG[ee];
            Bq[u] = yB[ee];
        }
#pragma unroll
        for (int u = 0; u < kPowerChunk; ++u) {
            Vq[u] = V[jj[u]];
            tq[u] = th[jj[u]];
        }
        const int cnt = min(kPowerChunk, e1 - e);
        for (int u = 0; u < cnt; ++u) {
            double s, c;
            sincos(ti - tq[0], &s, &c);
            const double vv = vi * Vq[0];
            const double pe = vv * (G[0] * c + Bq[0] * s);
            const double qe = vv * (G[0] * s - Bq[0] * c);
            Pij[e + u] = pe;
            Qij[e + u] = qe;
            p += pe;
            q += qe;
#pragma unroll
            for (int k = 0; k + 1 < kPowerChunk; ++k) {
                G[k] = G[k + 1];
                Bq[k] = Bq[k + 1];
                Vq[k] = Vq[k + 1];
                tq[k] = tq[k + 1];
            }
        }
    }
    P[i] = p;
    Q[i] = q;
}

// Flows at both ends of every branch (results; the solver loop evaluates the
// measured ones inside k_meas).
__global__ __launch_bounds__(kBlock) void k_flows(int nb, const int32_t* __restrict__ from,
                                                  const int32_t* __restrict__ to, const double* __restrict__ gs,
                                                  const double* __restrict__ bs, const double* __restrict__ bc,
                                                  const double* __restrict__ tap, const double* __restrict__ shift,
                                                  const uint8_t* __restrict__ status, const double* __restrict__ V,
                                                  const double* __restrict__ th, double* __restrict__ Pf,
                                                  double* __restrict__ Qf, double* __restrict__ Pt,
                                                  double* __restrict__ Qt) {
    int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= nb) return;
    if (!status[k]) {
        Pf[k] = Qf[k] = Pt[k] = Qt[k] = 0.0;
        return;
    }
    int i = from[k], j = to[k];
    double g = gs[k], b = bs[k], a = tap[k], bsh = b + 0.5 * bc[k];
    double s, c;
    sincos(th[i] - th[j] - shift[k], &s, &c);
    double Vi = V[i], Vj = V[j], A = Vi * Vj / a;
    Pf[k] = Vi * Vi * g / (a * a) - A * (g * c + b * s);
    Qf[k] = -Vi * Vi * bsh / (a * a) - A * (g * s - b * c);
    Pt[k] = Vj * Vj * g - A * (g * c - b * s);
    Qt[k] = -Vj * Vj * bsh + A * (g * s + b * c);
}

// ---------------------------------------------------------------------------
// Fused measurement pass: h(x), r = z - h, Huber weight, sqrt(w) r, the
// Jacobian row scaled by sqrt(w), and block partials of J and max |r|/sigma.
// ---------------------------------------------------------------------------
struct MeasArgs {
    int m, n;
    int robust, unit_weights, want_jac;
    int newton;  // Huber near the solution: outliers' gain weight (gamma/|u|)^6 / sigma^2 (not in the RHS)
    const int32_t* pth;  // angle state of each bus (-1 for an angle reference)
    double gamma;
    const uint8_t* mcode;
    const int32_t* mloc;
    const double* z;
    const double* sigma;
    const int32_t* hrow;
    const int32_t* yrow;
    const int32_t* ycol;
    const double* yG;
    const double* yB;
    const double* Pij;
    const double* Qij;
    const double* P;
    const double* Q;
    const double* V;
    const double* th;
    const int32_t* from;
    const int32_t* to;
    const double* gs;
    const double* bs;
    const double* bc;
    const double* tap;
    const double* shift;
    double* rs;
    double* hr;            // optional: (sqrt(w) H)_q (sqrt(w) r)_i at the H^T position of q
    const int32_t* hpos;   // H^T position of each H entry
    double* hout;
    double* rout;
    double* jpart;
    double* rpart;
    int* ipart;
};

constexpr int kMeasChunk = 4;  // Ybus entries of an injection row loaded together

__global__ __launch_bounds__(kBlock) void k_meas(MeasArgs A, double* __restrict__ hs) {
    const int blk = blockIdx.x;
    const int i = blk * kBlock + threadIdx.x;
    double swg = 0.0, wr = 0.0;
    // one Jacobian entry d: sqrt(w_gain) d into the row-major Jacobian (for the
    // gain) and/or d w r into the H^T-ordered right-hand-side products
    auto put = [&](int q, double d) {
        if (A.want_jac) hs[q] = swg * d;
        if (A.hr) A.hr[A.hpos[q]] = d * wr;
    };
    double jw = 0.0, rn = -1.0;
    int ri = -1;
    if (i < A.m) {
        const uint8_t code = A.mcode[i];
        const int t = code & 15, end = (code >> 4) & 1;
        const bool act = (code >> 5) & 1;
        const int l = A.mloc[i];
        const int q0 = A.hrow[i], q1 = A.hrow[i + 1];
        double hv = 0.0;
        // flow partials wrt (th_from, th_to, V_from, V_to)
        double d[4] = {0, 0, 0, 0};
        int fb = 0, tb = 0;
        if (t == 0) {
            hv = A.V[l];
        } else if (t == 1) {
            hv = A.th[l];
        } else if (t == 2 || t == 3 || t == 7 || t == 8) {
            hv = (t == 2 || t == 7) ? A.P[l] : A.Q[l];
        } else {
            fb = A.from[l];
            tb = A.to[l];
            double g = A.gs[l], b = A.bs[l], a = A.tap[l], bsh = b + 0.5 * A.bc[l];
            double s, c;
            sincos(A.th[fb] - A.th[tb] - A.shift[l], &s, &c);
            double Vi = A.V[fb], Vj = A.V[tb], AA = Vi * Vj / a;
            double p, q, dP[4], dQ[4];
            if (end == 0) {
                p = Vi * Vi * g / (a * a) - AA * (g * c + b * s);
                q = -Vi * Vi * bsh / (a * a) - AA * (g * s - b * c);
                dP[0] = AA * (g * s - b * c);
                dP[1] = -dP[0];
                dP[2] = 2.0 * Vi * g / (a * a) - (Vj / a) * (g * c + b * s);
                dP[3] = -(Vi / a) * (g * c + b * s);
                dQ[0] = -AA * (g * c + b * s);
                dQ[1] = -dQ[0];
                dQ[2] = -2.0 * Vi * bsh / (a * a) - (Vj / a) * (g * s - b * c);
                dQ[3] = -(Vi / a) * (g * s - b * c);
            } else {
                p = Vj * Vj * g - AA * (g * c - b * s);
                q = -Vj * Vj * bsh + AA * (g * s + b * c);
                dP[0] = AA * (g * s + b * c);
                dP[1] = -dP[0];
                dP[2] = -(Vj / a) * (g * c - b * s);
                dP[3] = 2.0 * Vj * g - (Vi / a) * (g * c - b * s);
                dQ[0] = AA * (g * c - b * s);
                dQ[1] = -dQ[0];
                dQ[2] = (Vj / a) * (g * s + b * c);
                dQ[3] = -2.0 * Vj * bsh + (Vi / a) * (g * s + b * c);
            }
            if (t == 4) {
                hv = p;
#pragma unroll
                for (int k = 0; k < 4; ++k) d[k] = dP[k];
            } else if (t == 5) {
                hv = q;
#pragma unroll
                for (int k = 0; k < 4; ++k) d[k] = dQ[k];
            } else {  // I_MAG = |S| / V_end
                double ve = end ? Vj : Vi;
                double S = sqrt(p * p + q * q);
                hv = S / ve;
                double inv = S > 1e-12 ? 1.0 / (S * ve) : 0.0;
#pragma unroll
                for (int k = 0; k < 4; ++k) d[k] = (p * dP[k] + q * dQ[k]) * inv;
                d[end ? 3 : 2] -= S / (ve * ve);
            }
        }
        const double r = act ? A.z[i] - hv : 0.0;
        const double sg = A.sigma[i];
        const double wb = 1.0 / (sg * sg);
        double w = wb;
        bool outlier = false;
        if (A.robust) {
            double ar = fabs(r);
            if (ar / sg > A.gamma) {
                w = A.gamma / (sg * ar);
                outlier = true;
            }
        }
        if (!act) w = 0.0;
        const double unit = act ? 1.0 : 0.0;
        const double sw = A.unit_weights ? unit : sqrt(w);
        // gain weight: IRLS, or near the solution (gamma/|u|)^6 / sigma^2 for the
        // outliers (|u| = |r| / sigma > gamma)
        const double xg = A.gamma * sg / fabs(r);
        swg = A.unit_weights ? unit : (A.newton && outlier ? xg * xg * xg / sg : sw);
        wr = A.unit_weights ? unit * r : w * r;
        A.rs[i] = sw * r;
        if (A.hout) {
            A.hout[i] = act ? hv : 0.0;
            A.rout[i] = r;
        }
        if (act) {
            jw = wb * r * r;
            rn = fabs(r) / sg;
            ri = i;
        }
        if ((A.want_jac || A.hr) && q1 > q0) {
            if (t == 0 || t == 1) {
                put(q0, 1.0);
            } else if (t == 2 || t == 3 || t == 7 || t == 8) {
                const bool isP = (t == 2 || t == 7);
                const int e0 = A.yrow[l], e1 = A.yrow[l + 1];
                const int nth = (q1 - q0) - (e1 - e0);
                int kth = 0;
                const double Vi = A.V[l], Pl = A.P[l], Ql = A.Q[l];
                // the row's entries in chunks of kMeasChunk: a chunk's loads are
                // issued together (indices clamped to the row's last entry)
                for (int e = e0; e < e1; e += kMeasChunk) {
                    int jj[kMeasChunk], pj[kMeasChunk];
                    double ga[kMeasChunk], ba[kMeasChunk], pa[kMeasChunk], qa[kMeasChunk], va[kMeasChunk];
#pragma unroll
                    for (int u = 0; u < kMeasChunk; ++u) {
                        const int ee = min(e + u, e1 - 1);
                        jj[u] = A.ycol[ee];
                        ga[u] = A.yG[ee];
                        ba[u] = A.yB[ee];
                        pa[u] = A.Pij[ee];
                        qa[u] = A.Qij[ee];
                    }
#pragma unroll
                    for (int u = 0; u < kMeasChunk; ++u) {
                        va[u] = A.V[jj[u]];
                        pj[u] = A.pth[jj[u]];
                    }
#pragma unroll
                    for (int u = 0; u < kMeasChunk; ++u) {
                        if (e + u >= e1) break;
                        double dth, dv;
                        if (jj[u] == l) {
                            if (isP) {
                                dth = -Ql - Vi * Vi * ba[u];
                                dv = Pl / Vi + Vi * ga[u];
                            } else {
                                dth = Pl - Vi * Vi * ga[u];
                                dv = Ql / Vi - Vi * ba[u];
                            }
                        } else if (isP) {
                            dth = qa[u];
                            dv = pa[u] / va[u];
                        } else {
                            dth = -pa[u];
                            dv = qa[u] / va[u];
                        }
                        if (pj[u] >= 0) put(q0 + kth++, dth);
                        put(q0 + nth + (e + u - e0), dv);
                    }
                }
            } else {
                // sorted columns: theta_min, theta_max, V_min, V_max (reference angles absent)
                int q = q0;
                if (fb == tb) {
                    if (A.pth[fb] >= 0) put(q++, d[0] + d[1]);
                    put(q, d[2] + d[3]);
                } else {
                    const bool fl = (code >> 6) & 1;
                    int lo = fl ? fb : tb, hi = fl ? tb : fb;
                    double dlo_t = fl ? d[0] : d[1], dhi_t = fl ? d[1] : d[0];
                    double dlo_v = fl ? d[2] : d[3], dhi_v = fl ? d[3] : d[2];
                    if (A.pth[lo] >= 0) put(q++, dlo_t);
                    if (A.pth[hi] >= 0) put(q++, dhi_t);
                    put(q++, dlo_v);
                    put(q, dhi_v);
                }
            }
        }
    }
    // block partials (deterministic: fixed tree order)
    __shared__ double sj[kBlock / kWave], sr[kBlock / kWave];
    __shared__ int si[kBlock / kWave];
    jw = wave_sum(jw);
    wave_argmax(rn, ri);
    const int w = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) {
        sj[w] = jw;
        sr[w] = rn;
        si[w] = ri;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double J = 0.0, R = -1.0;
        int I = -1;
        for (int k = 0; k < kBlock / kWave; ++k) {
            J += sj[k];
            if (sr[k] > R || (sr[k] == R && si[k] >= 0 && (I < 0 || si[k] < I))) {
                R = sr[k];
                I = si[k];
            }
        }
        A.jpart[blk] = J;
        A.rpart[blk] = R;
        A.ipart[blk] = I;
    }
}

// ---------------------------------------------------------------------------
// Right-hand side H^T W r by location, for the iterations on a cached factor
// (no Jacobian values needed).  Every H entry of an injection row is a value
// of its bus's Ybus row (k_power's per-entry terms), and every entry of a flow
// row is one of four partials of its branch end; so the rows of one location
// collapse to their summed weighted residuals, and each state gathers its RHS
// from the locations it touches:
//   k_meas_loc: one thread per bus (its V, theta and injection rows) and per
//               branch (the flow rows at both ends: one sincos per branch),
//               giving per bus the sums of w r by kind and per branch the
//               summed w r d for (theta_from, theta_to, V_from, V_to);
//   k_rhs_loc:  one thread per bus j, over its Ybus row (through the
//               transpose entry, the partial of the neighbour's injection)
//               and its incident branches, writing both of its states.
// The measurement rows must be grouped by location (device order with
// measurement classes): V/theta rows by bus, injection rows by bus, flow rows
// by branch end.  Same weights as k_meas (IRLS / Huber), different summation
// order.
// ---------------------------------------------------------------------------
struct LocArgs {
    int n, nb;
    int robust;
    double gamma;
    const uint8_t* mcode;
    const double* z;
    const double* sigma;
    const int32_t* bv_ptr;  // per device bus: its V / theta rows
    const int32_t* bi_ptr;  // per device bus: its injection rows
    const int32_t* br_ptr;  // per device branch end (2k + end): its flow rows
    const double* V;
    const double* th;
    const int32_t* from;
    const int32_t* to;
    const double* gs;
    const double* bs;
    const double* bc;
    const double* tap;
    const double* shift;
    double* sbus;  // per bus: sum of w r over its V, theta, P, Q rows
    double* sbr;   // per branch: sum of w r d over its flow rows, for theta_f, theta_t, V_f, V_t
    // the injections, computed by the bus threads first (k_power's work, fused)
    const int32_t* yrow;
    const int32_t* ycol;
    const double* yG;
    const double* yB;
    double* Pout;
    double* Qout;
    DevStatus* st;
    int reset;  // clear the iteration's status flags (the pass that opens an iteration)
    // optional: per block, J (base weights) and the largest |r| / sigma with its row
    double* jpart;
    double* rpart;
    int* ipart;
};

// w r of one row from its code, value and sigma at the evaluated h; jt gets the
// row's term of J, r^2 / sigma^2 with the base weight (0 for inactive rows)
struct LocWeight {
    int robust;
    double gamma;
    __device__ double operator()(uint8_t code, double z, double sg, double hv, double& jt) const {
        jt = 0.0;
        if (!((code >> 5) & 1)) return 0.0;  // inactive: weight 0
        const double r = z - hv;
        const double wb = 1.0 / (sg * sg);
        double w = wb;
        if (robust) {
            const double ar = fabs(r);
            if (ar / sg > gamma) w = gamma / (sg * ar);
        }
        jt = wb * r * r;
        return w * r;
    }
    __device__ double operator()(uint8_t code, double z, double sg, double hv) const {
        double jt;
        return (*this)(code, z, sg, hv, jt);
    }
};

// Entries per chunk of a thread's Ybus-row and measurement-row walks: the loads
// of a chunk are issued together (indices clamped to the last entry), so a walk
// waits for memory once per chunk rather than once per entry.
constexpr int kLocChunk = 4;
constexpr int kLocChunkFlow = 1;  // flow rows of one branch end, one at a time (fewer VGPRs: 120 -> 96)

template <int K = kLocChunk, typename F>
__device__ inline void loc_rows(const uint8_t* __restrict__ mcode, const double* __restrict__ zv,
                                const double* __restrict__ sigma, int i0, int i1, F f) {
    for (int i = i0; i < i1; i += K) {
        uint8_t c[K];
        double z[K], sg[K];
#pragma unroll
        for (int q = 0; q < K; ++q) {
            const int ii = min(i + q, i1 - 1);
            c[q] = mcode[ii];
            z[q] = zv[ii];
            sg[q] = sigma[ii];
        }
#pragma unroll
        for (int q = 0; q < K; ++q)
            if (i + q < i1) f(c[q], z[q], sg[q], i + q);
    }
}

// J with the base weights and the largest |r| / sigma (ties to the lower row)
// over the active rows a thread evaluates, from each row's J term r^2 / sigma^2
// (its square root is |r| / sigma: one root per thread, at the end)
struct LocStats {
    double J = 0.0, best = -1.0;
    int ri = -1;
    __device__ void add(uint8_t code, double jt, int i) {
        if (!((code >> 5) & 1)) return;
        J += jt;
        if (jt > best || (jt == best && (ri < 0 || i < ri))) {
            best = jt;
            ri = i;
        }
    }
    __device__ double rn() const { return best < 0.0 ? -1.0 : sqrt(best); }
};

// Block partials of LocStats (fixed tree order: deterministic), one per block.
__device__ inline void loc_stats_block(LocStats s, double* jpart, double* rpart, int* ipart) {
    __shared__ double sj[kBlock / kWave], sr[kBlock / kWave];
    __shared__ int si[kBlock / kWave];
    double J = wave_sum(s.J), R = s.rn();
    int I = s.ri;
    wave_argmax(R, I);
    const int w = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) {
        sj[w] = J;
        sr[w] = R;
        si[w] = I;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double JJ = 0.0, RR = -1.0;
        int II = -1;
        for (int k = 0; k < kBlock / kWave; ++k) {
            JJ += sj[k];
            if (sr[k] > RR || (sr[k] == RR && si[k] >= 0 && (II < 0 || si[k] < II))) {
                RR = sr[k];
                II = si[k];
            }
        }
        jpart[blockIdx.x] = JJ;
        rpart[blockIdx.x] = RR;
        ipart[blockIdx.x] = II;
    }
}

__global__ __launch_bounds__(kBlock) void k_meas_loc(LocArgs A) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    const LocWeight wr_of{A.robust, A.gamma};
    const uint8_t* mcode = A.mcode;
    const double* zv = A.z;
    const double* sigma = A.sigma;
    if (t == 0 && A.reset) {  // the iteration's status starts here (as in k_power)
        A.st->nan = 0;
        A.st->piv = 0;
        A.st->flow_err = 0;
        A.st->stop = 0;
    }
    const bool stats = A.jpart != nullptr;
    LocStats ls;
    if (t < A.n) {
        const int b = t;
        // injections of bus b over its Ybus row
        const double vi = A.V[b], ti = A.th[b];
        double pb = 0.0, qb = 0.0;
        const int e1 = A.yrow[b + 1];  // every row holds its diagonal entry
        for (int e = A.yrow[b]; e < e1; e += kLocChunk) {
            int jj[kLocChunk];
            double G[kLocChunk], Bq[kLocChunk], Vq[kLocChunk], tq[kLocChunk];
#pragma unroll
            for (int q = 0; q < kLocChunk; ++q) {
                const int ee = min(e + q, e1 - 1);
                jj[q] = A.ycol[ee];
                G[q] = A.yG[ee];
                Bq[q] = A.yB[ee];
            }
#pragma unroll
            for (int q = 0; q < kLocChunk; ++q) {
                Vq[q] = A.V[jj[q]];
                tq[q] = A.th[jj[q]];
            }
            // one sincos in the loop body: the chunk shifts down a register per entry
            const int cnt = min(kLocChunk, e1 - e);
            for (int q = 0; q < cnt; ++q) {
                double sn, cs;
                sincos(ti - tq[0], &sn, &cs);
                const double vv = vi * Vq[0];
                pb += vv * (G[0] * cs + Bq[0] * sn);
                qb += vv * (G[0] * sn - Bq[0] * cs);
#pragma unroll
                for (int k = 0; k + 1 < kLocChunk; ++k) {
                    G[k] = G[k + 1];
                    Bq[k] = Bq[k + 1];
                    Vq[k] = Vq[k + 1];
                    tq[k] = tq[k + 1];
                }
            }
        }
        A.Pout[b] = pb;
        A.Qout[b] = qb;
        double sV = 0.0, sT = 0.0, sP = 0.0, sQ = 0.0;
        // (selects rather than branches: a branch per sum lets the compiler
        // address the sums through a pointer, which moves them to scratch)
        loc_rows(mcode, zv, sigma, A.bv_ptr[b], A.bv_ptr[b + 1], [&](uint8_t c, double z, double sg, int i) {
            const bool isv = (c & 15) == 0;
            double jt;
            const double wr = wr_of(c, z, sg, isv ? vi : ti, jt);
            sV += isv ? wr : 0.0;
            sT += isv ? 0.0 : wr;
            if (stats) ls.add(c, jt, i);
        });
        loc_rows(mcode, zv, sigma, A.bi_ptr[b], A.bi_ptr[b + 1], [&](uint8_t c, double z, double sg, int i) {
            const int ty = c & 15;
            const bool isp = ty == 2 || ty == 7;
            double jt;
            const double wr = wr_of(c, z, sg, isp ? pb : qb, jt);
            sP += isp ? wr : 0.0;
            sQ += isp ? 0.0 : wr;
            if (stats) ls.add(c, jt, i);
        });
        double* o = A.sbus + 4 * (int64_t)b;
        o[0] = sV;
        o[1] = sT;
        o[2] = sP;
        o[3] = sQ;
    } else if (t < A.n + A.nb) {
        const int l = t - A.n;
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        if (A.br_ptr[2 * l + 2] > A.br_ptr[2 * l]) {
            const int fb = A.from[l], tb = A.to[l];
            const double g = A.gs[l], b = A.bs[l], a = A.tap[l], bsh = b + 0.5 * A.bc[l];
            double s, c;
            sincos(A.th[fb] - A.th[tb] - A.shift[l], &s, &c);
            const double Vi = A.V[fb], Vj = A.V[tb], AA = Vi * Vj / a;
            for (int end = 0; end < 2; ++end) {
                const int i0 = A.br_ptr[2 * l + end], i1 = A.br_ptr[2 * l + end + 1];
                if (i1 == i0) continue;
                // the flow at this end; its partials enter only through the row sums
                // cP = sum of w r over P rows (and I rows times p / (|S| V)), cQ
                // alike, cV = the I rows' extra V_end term
                const double cgb = end ? g * c - b * s : g * c + b * s;
                const double sgb = end ? g * s + b * c : g * s - b * c;
                const double p = end ? Vj * Vj * g - AA * cgb : Vi * Vi * g / (a * a) - AA * cgb;
                const double q = end ? -Vj * Vj * bsh + AA * sgb : -Vi * Vi * bsh / (a * a) - AA * sgb;
                const double ve = end ? Vj : Vi;
                double cP = 0.0, cQ = 0.0, cV = 0.0;
                loc_rows<kLocChunkFlow>(mcode, zv, sigma, i0, i1, [&](uint8_t cd, double z, double sg, int i) {
                    const int ty = cd & 15;
                    double jt;
                    if (ty == 4 || ty == 5) {
                        const double wr = wr_of(cd, z, sg, ty == 4 ? p : q, jt);
                        cP += ty == 4 ? wr : 0.0;
                        cQ += ty == 4 ? 0.0 : wr;
                    } else {  // I_MAG = |S| / V_end
                        const double S = sqrt(p * p + q * q);
                        const double inv = S > 1e-12 ? 1.0 / (S * ve) : 0.0;
                        const double wr = wr_of(cd, z, sg, S / ve, jt);
                        cP += wr * p * inv;
                        cQ += wr * q * inv;
                        cV -= wr * S / (ve * ve);
                    }
                    if (stats) ls.add(cd, jt, i);
                });
                // partials wrt (theta_f, theta_t, V_f, V_t)
                double dP[4], dQ[4];
                if (end == 0) {
                    dP[0] = AA * sgb;
                    dP[2] = 2.0 * Vi * g / (a * a) - (Vj / a) * cgb;
                    dP[3] = -(Vi / a) * cgb;
                    dQ[0] = -AA * cgb;
                    dQ[2] = -2.0 * Vi * bsh / (a * a) - (Vj / a) * sgb;
                    dQ[3] = -(Vi / a) * sgb;
                } else {
                    dP[0] = AA * sgb;
                    dP[2] = -(Vj / a) * cgb;
                    dP[3] = 2.0 * Vj * g - (Vi / a) * cgb;
                    dQ[0] = AA * cgb;
                    dQ[2] = (Vj / a) * sgb;
                    dQ[3] = -2.0 * Vj * bsh + (Vi / a) * sgb;
                }
                dP[1] = -dP[0];
                dQ[1] = -dQ[0];
#pragma unroll
                for (int k = 0; k < 4; ++k) acc[k] += cP * dP[k] + cQ * dQ[k];
                if (end)  // constant indices keep acc in registers
                    acc[3] += cV;
                else
                    acc[2] += cV;
            }
        }
        double* o = A.sbr + 4 * (int64_t)l;
#pragma unroll
        for (int k = 0; k < 4; ++k) o[k] = acc[k];
    }
    if (stats) loc_stats_block(ls, A.jpart, A.rpart, A.ipart);
}

template <typename TO>
__global__ __launch_bounds__(kBlock) void k_rhs_loc(int n, const int32_t* __restrict__ yrow,
                                                    const int32_t* __restrict__ ycol, const double* __restrict__ yGt,
                                                    const double* __restrict__ yBt, const double* __restrict__ P,
                                                    const double* __restrict__ Q, const double* __restrict__ V,
                                                    const double* __restrict__ th, const double* __restrict__ sbus,
                                                    const double* __restrict__ sbr, const int32_t* __restrict__ inc_ptr,
                                                    const int32_t* __restrict__ inc, const int32_t* __restrict__ pth,
                                                    const int32_t* __restrict__ pv, TO* __restrict__ x,
                                                    DevStatus* st) {
    const int j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= n) return;
    if (j == 0) {  // opens an iteration whose measurement pass closed the previous one
        st->nan = 0;
        st->piv = 0;
        st->flow_err = 0;
    }
    const double Vj = V[j], tj = th[j], Pj = P[j], Qj = Q[j];
    double at = sbus[4 * (int64_t)j + 1], av = sbus[4 * (int64_t)j];
    // The Ybus row in chunks (see kLocChunk).  Entry (j, l) needs the terms of
    // entry (l, j) of bus l's injection, V_l V_j (G_lj cos + B_lj sin)(th_l -
    // th_j) and its Q twin: recomputed here from V_l, th_l and the transposed
    // admittances, by the same operations as k_meas_loc's row of l (so the
    // values are the ones that summed into P_l, Q_l).  The diagonal's own terms
    // are Vj^2 G and -Vj^2 B, the self-admittance parts of the partials.
    const int e1 = yrow[j + 1];
    for (int e = yrow[j]; e < e1; e += kLocChunk) {
        int l[kLocChunk];
        double G[kLocChunk], Bq[kLocChunk];
#pragma unroll
        for (int q = 0; q < kLocChunk; ++q) {
            const int ee = min(e + q, e1 - 1);
            l[q] = ycol[ee];
            G[q] = yGt[ee];
            Bq[q] = yBt[ee];
        }
        double sp[kLocChunk], sq[kLocChunk], Vl[kLocChunk], tl[kLocChunk];
#pragma unroll
        for (int q = 0; q < kLocChunk; ++q) {
            sp[q] = sbus[4 * (int64_t)l[q] + 2];
            sq[q] = sbus[4 * (int64_t)l[q] + 3];
            Vl[q] = V[l[q]];
            tl[q] = th[l[q]];
        }
        // one sincos in the loop body: the chunk shifts down a register per entry
        const int cnt = min(kLocChunk, e1 - e);
        for (int q = 0; q < cnt; ++q) {
            double sn, cs;
            sincos(tl[0] - tj, &sn, &cs);
            const double vv = Vl[0] * Vj;
            const double pe = vv * (G[0] * cs + Bq[0] * sn), qe = vv * (G[0] * sn - Bq[0] * cs);
            if (l[0] == j) {
                at += sp[0] * (qe - Qj) + sq[0] * (Pj - pe);
                av += sp[0] * ((Pj + pe) / Vj) + sq[0] * ((Qj + qe) / Vj);
            } else {
                // partials of bus l's injections wrt this bus: the terms of entry (l, j)
                at += sp[0] * qe - sq[0] * pe;
                av += (sp[0] * pe + sq[0] * qe) / Vj;
            }
#pragma unroll
            for (int k = 0; k + 1 < kLocChunk; ++k) {
                l[k] = l[k + 1];
                G[k] = G[k + 1];
                Bq[k] = Bq[k + 1];
                sp[k] = sp[k + 1];
                sq[k] = sq[k + 1];
                Vl[k] = Vl[k + 1];
                tl[k] = tl[k + 1];
            }
        }
    }
    const int c1 = inc_ptr[j + 1];
    for (int c = inc_ptr[j]; c < c1; c += kLocChunk) {
        int code[kLocChunk];
#pragma unroll
        for (int q = 0; q < kLocChunk; ++q) code[q] = inc[min(c + q, c1 - 1)];
        double o0[kLocChunk], o1[kLocChunk];
#pragma unroll
        for (int q = 0; q < kLocChunk; ++q) {  // (theta, V) sums of this bus's end of the branch
            const double* o = sbr + 4 * (int64_t)(code[q] >> 1) + (code[q] & 1);
            o0[q] = o[0];
            o1[q] = o[2];
        }
#pragma unroll
        for (int q = 0; q < kLocChunk; ++q) {
            if (c + q >= c1) break;
            at += o0[q];
            av += o1[q];
        }
    }
    if (pth[j] >= 0) x[pth[j]] = (TO)at;
    x[pv[j]] = (TO)av;
}

// Single-block finish of the J / largest-residual partials into st (J, rmax,
// ridx), NT threads.  Loads go four partials per thread at a time, so that the
// block waits for memory once per four strides.
template <int NT>
__device__ inline void jfinish_block(int nparts, const double* __restrict__ jpart, const double* __restrict__ rpart,
                                     const int* __restrict__ ipart, DevStatus* st) {
    double J = 0.0, R = -1.0;
    int I = -1;
    constexpr int U = 4;
    for (int p0 = threadIdx.x; p0 < nparts; p0 += U * NT) {
        double jv[U], rv[U];
        int iv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int p = min(p0 + u * NT, nparts - 1);
            jv[u] = jpart[p];
            rv[u] = rpart[p];
            iv[u] = ipart[p];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (p0 + u * NT >= nparts) break;
            J += jv[u];
            if (rv[u] > R || (rv[u] == R && iv[u] >= 0 && (I < 0 || iv[u] < I))) {
                R = rv[u];
                I = iv[u];
            }
        }
    }
    __shared__ double sj[NT / kWave], sr[NT / kWave];
    __shared__ int si[NT / kWave];
    J = wave_sum(J);
    wave_argmax(R, I);
    const int w = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) {
        sj[w] = J;
        sr[w] = R;
        si[w] = I;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double JJ = 0.0, RR = -1.0;
        int II = -1;
        for (int k = 0; k < NT / kWave; ++k) {
            JJ += sj[k];
            if (sr[k] > RR || (sr[k] == RR && si[k] >= 0 && (II < 0 || si[k] < II))) {
                RR = sr[k];
                II = si[k];
            }
        }
        st->J = JJ;
        st->rmax = RR;
        st->ridx = II;
    }
}

__global__ __launch_bounds__(kBlock) void k_jfinish(int nparts, const double* __restrict__ jpart,
                                                    const double* __restrict__ rpart, const int* __restrict__ ipart,
                                                    DevStatus* st) {
    jfinish_block<kBlock>(nparts, jpart, rpart, ipart, st);
}

// ---------------------------------------------------------------------------
// Gain entries G_e = sum over the cached product list of (sqrt(w)H)_a (sqrt(w)H)_b
// ---------------------------------------------------------------------------
constexpr int kSpgemmChunk = 4;  // product pairs of a G entry loaded together

template <typename T>
__global__ __launch_bounds__(kBlock) void k_spgemm(int nG, const int32_t* __restrict__ gptr,
                                                   const int2* __restrict__ gpair, const double* __restrict__ hs,
                                                   T* __restrict__ Gval) {
    int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= nG) return;
    // the pairs in chunks of kSpgemmChunk: their loads issued together
    // (indices clamped to the entry's last pair), then summed in pair order
    const int p0 = gptr[e], p1 = gptr[e + 1];
    double acc = 0.0;
    for (int p = p0; p < p1; p += kSpgemmChunk) {
        int2 pr[kSpgemmChunk];
#pragma unroll
        for (int u = 0; u < kSpgemmChunk; ++u) pr[u] = gpair[min(p + u, p1 - 1)];
        double a[kSpgemmChunk], b[kSpgemmChunk];
#pragma unroll
        for (int u = 0; u < kSpgemmChunk; ++u) {
            a[u] = hs[pr[u].x];
            b[u] = hs[pr[u].y];
        }
#pragma unroll
        for (int u = 0; u < kSpgemmChunk; ++u)
            if (p + u < p1) acc += a[u] * b[u];
    }
    Gval[e] = (T)acc;  // FP64 products, rounded once to the factor precision
}

// ---------------------------------------------------------------------------
// Multifrontal supernodal Cholesky.  One workgroup per front; a front is the
// dense c x c (column-major) matrix of supernode s: its G columns plus the
// extend-add of its children's update blocks.  The first k columns are
// factored; the trailing (c-k)^2 block is the update passed to the parent.
// ---------------------------------------------------------------------------
// Per-supernode and per-child-edge descriptors of the factorization: what a
// front's wave needs before its loads, in one load each instead of a chain
// (supernode -> its G range and child list -> each child's sizes and offsets).
struct FrontDesc {
    int32_t c, k, first, g0, ng, p0, nch, pad;
};
struct ChildDesc {  // child edge p (in child_ptr order): the child's sizes and offsets
    int32_t cc, kc, pad0, pad1;
    int64_t foff, rmoff;
};

template <typename T>
struct FactArgs {
    const FrontDesc* fdesc;
    const ChildDesc* cdesc;
    const int32_t* sn_first;
    const int32_t* sn_k;
    const int32_t* sn_c;
    const int64_t* foff;
    const int32_t* child_ptr;
    const int32_t* child;
    const int64_t* rmoff;
    const int32_t* rmap;
    const int32_t* g_sn_ptr;
    const int32_t* gloc;
    const int32_t* gdiag;
    const T* Gval;
    T* F;
    T piv_tol;
    DevStatus* st;
};

// Who works on one front: the whole workgroup (Team<false>, NT threads) or a
// single wave inside a larger workgroup (Team<true>, 64 lanes) when several
// waves each process a different front.
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
template <bool WAVE>
struct Team;
template <>
struct Team<false> {
    static __device__ int id() { return threadIdx.x; }
    static __device__ void sync() { __syncthreads(); }
};
template <>
struct Team<true> {
    static __device__ int id() { return threadIdx.x & (kWave - 1); }
    static __device__ void sync() { wave_sync(); }
};

// A batch of a child's update block for the extend-add: GB positions (a, b)
// of the u x u square per lane, stepped without a division per position;
// ra < 0 marks a position outside the lower triangle.  Every load of the batch
// (relative map entries, update values) is issued before any use.
constexpr int kUpdBatch = 8;
struct UpdWalk {
    int u, a, b, da, db, cc, kc;
};
template <typename T>
__device__ inline void upd_load(UpdWalk& w, const T* Fc, const int32_t* rm, int* ra, int* rb, T* v) {
#pragma unroll
    for (int q = 0; q < kUpdBatch; ++q) {
        const bool ok = w.b < w.u && w.a >= w.b;
        ra[q] = ok ? rm[w.a] : -1;
        rb[q] = ok ? rm[w.b] : 0;
        v[q] = ok ? Fc[(int64_t)(w.kc + w.b) * w.cc + w.kc + w.a] : T(0);
        w.a += w.da;
        w.b += w.db;
        if (w.a >= w.u) {
            w.a -= w.u;
            ++w.b;
        }
    }
}

// One wave factors one front with its dependent global loads gathered into
// few round trips: the front's descriptor; then its children's descriptors,
// the G entries and the pivot thresholds, lane-parallel; then the children's
// update blocks, with PF the next child's loads issued before the current
// child's adds (it needs the registers of a second batch: off where a
// 1024-thread workgroup caps a wave at 128 VGPRs).  (The factor sweeps are
// bound by these round trips: fronts have at most a few dozen rows.)
template <typename T, bool PF = true>
__device__ inline void factor_front_wave(const FactArgs<T>& A, int s, T* F) {
    const int lane = threadIdx.x & (kWave - 1);
    const FrontDesc d = A.fdesc[s];
    const int c = d.c, k = d.k, first = d.first, g0 = d.g0, g1 = d.g0 + d.ng, p0 = d.p0, nch = d.nch;
    int cc_l = 0, kc_l = 0;
    int64_t fo_l = 0, ro_l = 0;
    if (lane < nch) {
        const ChildDesc cd = A.cdesc[p0 + lane];
        cc_l = cd.cc;
        kc_l = cd.kc;
        fo_l = cd.foff;
        ro_l = cd.rmoff;
    }
    const int ge_l = lane < k ? A.gdiag[first + lane] : -1;
    for (int idx = lane; idx < c * c; idx += kWave) F[idx] = T(0);
    wave_sync();
    // the G entries: a batch of loads per lane before their LDS stores (one
    // round trip per batch, not per entry)
    constexpr int GB = 8;
    for (int e0 = g0 + lane; e0 < g1; e0 += GB * kWave) {
        int32_t gl[GB];
        T gv[GB];
#pragma unroll
        for (int u = 0; u < GB; ++u) {
            const int e = e0 + u * kWave;
            gl[u] = e < g1 ? A.gloc[e] : 0;
            gv[u] = e < g1 ? A.Gval[e] : T(0);
        }
#pragma unroll
        for (int u = 0; u < GB; ++u)
            if (e0 + u * kWave < g1) F[gl[u]] = gv[u];
    }
    const T gdl = ge_l >= 0 ? A.Gval[ge_l] : T(0);
    wave_sync();
    // child p's walk over its update block, and the block's address
    auto child = [&](int p, UpdWalk& w, const T*& Fc, const int32_t*& rm) {
        int64_t fo, ro;
        if (p < kWave) {
            w.cc = __shfl(cc_l, p, kWave);
            w.kc = __shfl(kc_l, p, kWave);
            fo = __shfl(fo_l, p, kWave);
            ro = __shfl(ro_l, p, kWave);
        } else {  // more than 64 children: the rest one by one
            const ChildDesc cd = A.cdesc[p0 + p];
            w.cc = cd.cc;
            w.kc = cd.kc;
            fo = cd.foff;
            ro = cd.rmoff;
        }
        w.u = w.cc - w.kc;
        const int uu = max(w.u, 1);
        w.da = kWave % uu;
        w.db = kWave / uu;
        w.b = w.u > 0 ? lane / uu : 0;
        w.a = lane - w.b * uu;
        Fc = A.F + fo;
        rm = A.rmap + ro;
    };
    if (nch > 0) {
        UpdWalk wc, wn;
        const T *Fc, *Fn = nullptr;
        const int32_t *rmc, *rmn = nullptr;
        int ra[kUpdBatch], rb[kUpdBatch], ran[kUpdBatch], rbn[kUpdBatch];
        T v[kUpdBatch], vn[kUpdBatch];
        child(0, wc, Fc, rmc);
        upd_load<T>(wc, Fc, rmc, ra, rb, v);
        for (int p = 0; p < nch; ++p) {
            if (PF && p + 1 < nch) {  // the next child's first batch in flight during this child's adds
                child(p + 1, wn, Fn, rmn);
                upd_load<T>(wn, Fn, rmn, ran, rbn, vn);
            }
            for (;;) {  // positions of one child are distinct: no barrier between its batches
#pragma unroll
                for (int q = 0; q < kUpdBatch; ++q)
                    if (ra[q] >= 0) F[rb[q] * c + ra[q]] += v[q];
                if (wc.b >= wc.u) break;
                upd_load<T>(wc, Fc, rmc, ra, rb, v);
            }
            wave_sync();  // children may add into the same positions
            if (!PF && p + 1 < nch) {
                child(p + 1, wc, Fc, rmc);
                upd_load<T>(wc, Fc, rmc, ra, rb, v);
            } else if (p + 1 < nch) {
                wc = wn;
                Fc = Fn;
                rmc = rmn;
#pragma unroll
                for (int q = 0; q < kUpdBatch; ++q) {
                    ra[q] = ran[q];
                    rb[q] = rbn[q];
                    v[q] = vn[q];
                }
            }
        }
    }
    // The panel (columns 0..k-1), left-looking: column j from the columns
    // before it, one lane per row (rows j + lane, + 64, ...); the diagonal
    // comes from lane 0 by a shuffle, so a column costs one pass and one sync.
    for (int j = 0; j < k; ++j) {
        const T gd = j < kWave ? __shfl(gdl, j, kWave) : [&] {
            const int ge = A.gdiag[first + j];
            return ge >= 0 ? A.Gval[ge] : T(0);
        }();
        T dj = T(1);
        for (int r0 = j; r0 < c; r0 += kWave) {
            const int r = r0 + lane;
            const int rr = r < c ? r : j;
            T acc = F[j * c + rr];
            int m = 0;
            for (; m + 4 <= j; m += 4) {
                T x[4], y[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    x[q] = F[(m + q) * c + rr];
                    y[q] = F[(m + q) * c + j];
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) acc -= x[q] * y[q];
            }
            for (; m < j; ++m) acc -= F[m * c + rr] * F[m * c + j];
            if (r0 == j) {  // lane 0 holds the diagonal
                T dd = __shfl(acc, 0, kWave);
                if (!(dd > A.piv_tol * gd) || !(dd > T(0))) {
                    if (lane == 0) A.st->piv = 1;
                    dd = gd > T(0) ? A.piv_tol * gd : T(1);
                }
                dj = sqrt(dd);
            }
            if (r < c) F[j * c + r] = r == j ? dj : acc / dj;
        }
        wave_sync();
    }
    // the trailing block in one pass: F_ab -= sum_{m<k} L_am L_bm, k <= b <= a
    const int w = c - k;
    if (w > 0) {
        const int da = kWave % w, db = kWave / w;
        int b = lane / w, a = lane - b * w;
        while (b < w) {
            if (a >= b) {
                const int aa = k + a, bb = k + b;
                T acc = F[bb * c + aa];
                int m = 0;
                for (; m + 4 <= k; m += 4) {
                    T x[4], y[4];
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        x[q] = F[(m + q) * c + aa];
                        y[q] = F[(m + q) * c + bb];
                    }
#pragma unroll
                    for (int q = 0; q < 4; ++q) acc -= x[q] * y[q];
                }
                for (; m < k; ++m) acc -= F[m * c + aa] * F[m * c + bb];
                F[bb * c + aa] = acc;
            }
            a += da;
            b += db;
            if (a >= w) {
                a -= w;
                ++b;
            }
        }
    }
    wave_sync();
}

template <typename T, int NT, bool WAVE = false, bool PF = true>
__device__ inline void factor_front(const FactArgs<T>& A, int s, T* F) {
    if (NT == kWave) {
        factor_front_wave<T, PF>(A, s, F);
        return;
    }
    using Tm = Team<WAVE>;
    const int tid = Tm::id();
    const int c = A.sn_c[s], k = A.sn_k[s], first = A.sn_first[s];
    for (int idx = tid; idx < c * c; idx += NT) F[idx] = T(0);
    Tm::sync();
    for (int e = A.g_sn_ptr[s] + tid; e < A.g_sn_ptr[s + 1]; e += NT) F[A.gloc[e]] = A.Gval[e];
    Tm::sync();
    for (int p = A.child_ptr[s]; p < A.child_ptr[s + 1]; ++p) {
        const int ch = A.child[p];
        const int cc = A.sn_c[ch], kc = A.sn_k[ch], u = cc - kc;
        const T* Fc = A.F + A.foff[ch];
        const int32_t* rm = A.rmap + A.rmoff[ch];
        // the u x u square walked NT entries at a time; (a, b) advances by the
        // precomputed (NT % u, NT / u) instead of a division per entry
        if (u > 0) {
            const int da = NT % u, db = NT / u;
            int b = tid / u, a = tid - b * u;
            while (b < u) {
                if (a >= b) F[rm[b] * c + rm[a]] += Fc[(int64_t)(kc + b) * cc + kc + a];
                a += da;
                b += db;
                if (a >= u) {
                    a -= u;
                    ++b;
                }
            }
        }
        Tm::sync();
    }
    // the pivot thresholds' G_jj, loaded together before the pivot loop (one
    // lane per pivot) instead of one dependent load pair per pivot
    const bool pre = NT == kWave && k <= kWave;
    T gdl = T(0);
    if (pre && tid < k) {
        const int ge = A.gdiag[first + tid];
        gdl = ge >= 0 ? A.Gval[ge] : T(0);
    }
    for (int j = 0; j < k; ++j) {
        T dd = F[j * c + j];
        T gd;
        if (pre) {
            gd = __shfl(gdl, j, kWave);
        } else {
            const int ge = A.gdiag[first + j];
            gd = ge >= 0 ? A.Gval[ge] : T(0);
        }
        Tm::sync();
        if (!(dd > A.piv_tol * gd) || !(dd > T(0))) {
            // below threshold: flag it and substitute threshold * G_jj, so the
            // factor stays SPD (usable as a preconditioner)
            if (tid == 0) A.st->piv = 1;
            dd = gd > T(0) ? A.piv_tol * gd : T(1);
        }
        const T dj = sqrt(dd);
        if (tid == 0) F[j * c + j] = dj;
        for (int i = j + 1 + tid; i < c; i += NT) F[j * c + i] = F[j * c + i] / dj;
        Tm::sync();
        // rank-1 update of the trailing lower triangle over the w x w square,
        // NT entries at a time with the same division-free stepping
        const int w = c - j - 1;
        if (w > 0) {
            const int da = NT % w, db = NT / w;
            int b = tid / w, a = tid - b * w;
            while (b < w) {
                if (a >= b) {
                    const int aa = j + 1 + a, bb = j + 1 + b;
                    F[bb * c + aa] -= F[j * c + aa] * F[j * c + bb];
                }
                a += da;
                b += db;
                if (a >= w) {
                    a -= w;
                    ++b;
                }
            }
        }
        Tm::sync();
    }
}

template <typename T, int CM>
__global__ __launch_bounds__(kWave) void k_factor_lds(FactArgs<T> A, const int32_t* __restrict__ list) {
    __shared__ T F[CM * CM];
    const int s = list[blockIdx.x];
    factor_front<T, kWave>(A, s, F);
    const int c = A.sn_c[s];
    T* Fo = A.F + A.foff[s];
    for (int idx = threadIdx.x; idx < c * c; idx += kWave) Fo[idx] = F[idx];
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_factor_global(FactArgs<T> A, const int32_t* __restrict__ list) {
    const int s = list[blockIdx.x];
    factor_front<T, kBlock>(A, s, A.F + A.foff[s]);
}

// ---------------------------------------------------------------------------
// Triangular solves over the supernodal tree (one wave per front).
// Forward: f = [b_s ; 0] + extend-add of children's update vectors,
//          L11 y = f_top, u_s = f_bot - L21 y.
// Backward: x_s = L11^-T (y - L21^T x_rows).
// ---------------------------------------------------------------------------
template <typename T>
struct SolveArgs {
    const int32_t* sn_first;
    const int32_t* sn_k;
    const int32_t* sn_c;
    const int64_t* foff;
    const int64_t* rows_off;
    const int32_t* rows;
    const int32_t* child_ptr;
    const int32_t* child;
    const int64_t* rmoff;
    const int32_t* rmap;
    const int64_t* uoff;
    const T* F;
    T* x;
    T* ubuf;
    const T* M;            // per-front solve matrices [L11^-1; -L21 L11^-1] (c x k, column-major)
    const int64_t* moff;
};

template <typename T, bool WAVE = false>
__device__ inline void fwd_front(const SolveArgs<T>& A, int s, T* f) {
    using Tm = Team<WAVE>;
    const int tid = Tm::id();
    const int c = A.sn_c[s], k = A.sn_k[s], first = A.sn_first[s];
    const T* L = A.F + A.foff[s];
    for (int i = tid; i < c; i += kWave) f[i] = i < k ? A.x[first + i] : T(0);
    Tm::sync();
    for (int p = A.child_ptr[s]; p < A.child_ptr[s + 1]; ++p) {
        const int ch = A.child[p];
        const int u = A.sn_c[ch] - A.sn_k[ch];
        const T* uc = A.ubuf + A.uoff[ch];
        const int32_t* rm = A.rmap + A.rmoff[ch];
        for (int a = tid; a < u; a += kWave) f[rm[a]] += uc[a];
        Tm::sync();
    }
    for (int j = 0; j < k; ++j) {
        const T yj = f[j] / L[j * c + j];
        Tm::sync();
        if (tid == 0) f[j] = yj;
        for (int i = j + 1 + tid; i < c; i += kWave) f[i] -= L[j * c + i] * yj;
        Tm::sync();
    }
    for (int i = tid; i < k; i += kWave) A.x[first + i] = f[i];
    T* us = A.ubuf + A.uoff[s];
    for (int a = tid; a < c - k; a += kWave) us[a] = f[k + a];
}

template <typename T, bool WAVE = false>
__device__ inline void bwd_front(const SolveArgs<T>& A, int s, T* xs) {
    using Tm = Team<WAVE>;
    const int tid = Tm::id();
    const int c = A.sn_c[s], k = A.sn_k[s], first = A.sn_first[s];
    const T* L = A.F + A.foff[s];
    const int32_t* rows = A.rows + A.rows_off[s];
    for (int i = tid; i < c; i += kWave) xs[i] = i < k ? A.x[first + i] : A.x[rows[i]];
    Tm::sync();
    for (int j = k - 1; j >= 0; --j) {
        double part = 0.0;
        for (int i = j + 1 + tid; i < c; i += kWave) part += (double)(L[j * c + i] * xs[i]);
        part = wave_sum(part);
        if (tid == 0) xs[j] = (T)(((double)xs[j] - part) / (double)L[j * c + j]);
        Tm::sync();
    }
    for (int i = tid; i < k; i += kWave) A.x[first + i] = xs[i];
}

// Per-front solve matrix M = [W; -L21 W], W = L11^-1 (lower triangular), so
// that a front's forward step is f <- M f_top + [0; f_bot] and its backward
// step x_s = M^T [y_s; x_rows]: one parallel product instead of k dependent
// pivot steps.  One wave per front; lane jj builds column jj.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_front_m(int nsn, SolveArgs<T> A, T* __restrict__ M,
                                                    const int32_t* __restrict__ list) {
    const int q = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    if (q >= nsn) return;  // uniform per wave
    const int s = list ? list[q] : q;
    const int c = A.sn_c[s], k = A.sn_k[s];
    const T* L = A.F + A.foff[s];
    T* Ms = M + A.moff[s];
    for (int jj = lane; jj < k; jj += kWave) {
        T* w = Ms + (int64_t)jj * c;
        for (int i = 0; i < jj; ++i) w[i] = T(0);
        for (int i = jj; i < k; ++i) {
            double acc = i == jj ? 1.0 : 0.0;
            for (int j = jj; j < i; ++j) acc -= (double)L[j * c + i] * (double)w[j];
            w[i] = (T)(acc / (double)L[i * c + i]);
        }
        for (int i = k; i < c; ++i) {
            double acc = 0.0;
            for (int j = jj; j < k; ++j) acc -= (double)L[j * c + i] * (double)w[j];
            w[i] = (T)acc;
        }
    }
}

template <typename T, int VM>
__global__ __launch_bounds__(kWave) void k_fwd(SolveArgs<T> A, const int32_t* __restrict__ list) {
    __shared__ T f[VM];
    fwd_front<T>(A, list[blockIdx.x], f);
}

template <typename T, int VM>
__global__ __launch_bounds__(kWave) void k_bwd(SolveArgs<T> A, const int32_t* __restrict__ list) {
    __shared__ T xs[VM];
    bwd_front<T>(A, list[blockIdx.x], xs);
}

// ---------------------------------------------------------------------------
// Task sweeps: one workgroup walks a whole group of fronts level by level
// (fronts of one level spread over its waves, one front per wave at a time,
// __syncthreads between levels).  A task is either a complete bottom subtree
// of the supernodal tree or the narrow top of the tree; all data exchanged
// between its fronts stays inside one CU, so no cross-XCD ordering is needed
// below the kernel boundary.  MODE 0 factor, 1 forward solve, 2 backward solve
// (levels walked top-down).
// ---------------------------------------------------------------------------
constexpr int kBottomThreads = 256;  // workgroup of a bottom factor task

template <typename T, int MODE, int MAXT = 1024>
__global__ __launch_bounds__(MAXT) void k_sweep_tasks(FactArgs<T> FA, SolveArgs<T> SA,
                                                      const int32_t* __restrict__ fronts,
                                                      const int32_t* __restrict__ lvl_ptr,
                                                      const int32_t* __restrict__ lvl_off, int task_base,
                                                      int slice) {
    extern __shared__ unsigned char smem_raw[];
    const int task = task_base + blockIdx.x;
    const int nw = blockDim.x / kWave, w = threadIdx.x / kWave;
    T* my = reinterpret_cast<T*>(smem_raw) + (size_t)w * slice;
    const int l0 = lvl_off[task];
    const int nlev = lvl_off[task + 1] - l0 - 1;
    for (int li = 0; li < nlev; ++li) {
        const int L = MODE == 2 ? nlev - 1 - li : li;
        const int a = lvl_ptr[l0 + L], b = lvl_ptr[l0 + L + 1];
        for (int q = a + w; q < b; q += nw) {
            const int s = fronts[q];
            if (MODE == 0) {
                factor_front<T, kWave, true, (MAXT <= 512)>(FA, s, my);
                const int c = FA.sn_c[s];
                T* Fo = FA.F + FA.foff[s];
                for (int idx = threadIdx.x & (kWave - 1); idx < c * c; idx += kWave) Fo[idx] = my[idx];
            } else if (MODE == 1) {
                fwd_front<T, true>(SA, s, my);
            } else {
                bwd_front<T, true>(SA, s, my);
            }
            wave_sync();
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// LDS-resident task triangular solves.  A task (a bottom subtree, or a chunk of
// consecutive levels of the narrow top of the tree) is solved by one
// workgroup: it first loads the L panels of all its fronts, their metadata,
// relative maps and child lists into LDS with coalesced loads, then walks its
// levels with LDS-only traffic (one wave per front, __syncthreads per level),
// and finally writes back the solution of its columns and, for fronts whose
// parent lies outside the task, the update vector.  Per level this costs LDS
// latency instead of a chain of dependent HBM/L2 loads.
// ---------------------------------------------------------------------------
constexpr int kTH = 12;  // task header ints
enum { TH_NF = 0, TH_NLEV, TH_META, TH_LVL, TH_CH, TH_NCH, TH_RM, TH_NRM, TH_SRC, TH_NSRC, TH_SUML, TH_SUMF };
constexpr int kFM = 12;  // per-front ints
enum { FM_S = 0, FM_C, FM_K, FM_FIRST, FM_L, FM_F, FM_NCH, FM_CH, FM_RM, FM_EXTPAR, FM_SRC, FM_PAR };
constexpr int kLdsSpinLimit = 1 << 22;  // polls of an in-workgroup dependency before it is reported

struct LdsTaskArgs {
    const int32_t* th;
    const int32_t* meta;
    const int32_t* lvl;
    const int32_t* ch;
    const int32_t* rm;
    const int32_t* src;
    DevStatus* st;
};

// In-workgroup dependencies between the fronts of one task (FLAGS variants):
// a wave walks its fronts in dependency order and waits on LDS flags only for
// the fronts it needs, instead of a workgroup barrier per level.  Fronts are
// dealt to the waves round-robin in an order that lists every front after the
// ones it waits for, so the lowest waiting front can always proceed; the spin
// limit only guards against a plan that breaks that order (the sweep is then
// redone level by level).
__device__ inline void lds_wait(const int* flag, DevStatus* st) {
    int it = 0;
    while (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) == 0) {
        if (++it > kLdsSpinLimit) {
            st->flow_err = 1;
            break;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

__device__ inline void lds_publish(int* flag, int lane) {
    wave_sync();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    if (lane == 0) __hip_atomic_store(flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

template <typename T>
struct LdsView {
    T* Lp;
    T* fv;
    int* mt;
    int* lv;
    int* chl;
    int* rml;
    int* srl;
    int* flg;  // per-front completion flags (FLAGS variants)
    int nf, nlev;
};

template <typename T>
__device__ inline LdsView<T> lds_task_load(const LdsTaskArgs& A, const int32_t* H, unsigned char* smem, bool with_src) {
    LdsView<T> v;
    v.nf = H[TH_NF];
    v.nlev = H[TH_NLEV];
    v.Lp = reinterpret_cast<T*>(smem);
    v.fv = v.Lp + H[TH_SUML];
    v.mt = reinterpret_cast<int*>(v.fv + H[TH_SUMF]);
    v.lv = v.mt + v.nf * kFM;
    v.chl = v.lv + v.nlev + 1;
    v.rml = v.chl + H[TH_NCH];
    v.srl = v.rml + H[TH_NRM];
    v.flg = v.srl + H[TH_NSRC];
    const int nt = blockDim.x;
    for (int i = threadIdx.x; i < v.nf; i += nt) v.flg[i] = 0;
    for (int i = threadIdx.x; i < v.nf * kFM; i += nt) v.mt[i] = A.meta[H[TH_META] + i];
    for (int i = threadIdx.x; i <= v.nlev; i += nt) v.lv[i] = A.lvl[H[TH_LVL] + i];
    for (int i = threadIdx.x; i < H[TH_NCH]; i += nt) v.chl[i] = A.ch[H[TH_CH] + i];
    for (int i = threadIdx.x; i < H[TH_NRM]; i += nt) v.rml[i] = A.rm[H[TH_RM] + i];
    if (with_src)
        for (int i = threadIdx.x; i < H[TH_NSRC]; i += nt) v.srl[i] = A.src[H[TH_SRC] + i];
    return v;
}

template <typename T, bool FLAGS>
__device__ inline void fwd_lds_front(const LdsView<T>& v, int lf, int lane) {
    const int* m = v.mt + lf * kFM;
    const int c = m[FM_C], k = m[FM_K];
    T* f = v.fv + m[FM_F];
    const T* Lp = v.Lp + m[FM_L];
    for (int q = m[FM_CH]; q < m[FM_CH] + m[FM_NCH]; ++q) {
        const int lc = v.chl[q];
        if (lc < 0) continue;
        const int* mc = v.mt + lc * kFM;
        const int u = mc[FM_C] - mc[FM_K];
        const T* uc = v.fv + mc[FM_F] + mc[FM_K];
        const int* rmc = v.rml + mc[FM_RM];
        for (int a = lane; a < u; a += kWave) f[rmc[a]] += uc[a];
        wave_sync();
    }
    // f <- M f_top + [0; f_bot]  (c <= 64: one row per lane)
    T out = T(0);
    if (lane < c) {
        out = lane < k ? T(0) : f[lane];
        for (int j = 0; j < k; ++j) out += Lp[j * c + lane] * f[j];
    }
    wave_sync();
    if (lane < c) f[lane] = out;
    wave_sync();
}

template <typename T, bool FLAGS>
__global__ __launch_bounds__(1024) void k_fwd_lds(LdsTaskArgs A, SolveArgs<T> SA, int task_base) {
    extern __shared__ unsigned char smem_raw[];
    const int32_t* H = A.th + (size_t)(task_base + blockIdx.x) * kTH;
    LdsView<T> v = lds_task_load<T>(A, H, smem_raw, false);
    const int nw = blockDim.x / kWave, w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
    __syncthreads();
    // solve matrices (c x k per front, contiguous) and right-hand sides
    for (int lf = w; lf < v.nf; lf += nw) {
        const int* m = v.mt + lf * kFM;
        const int s = m[FM_S], c = m[FM_C], k = m[FM_K], first = m[FM_FIRST];
        const T* Mg = SA.M + SA.moff[s];
        for (int e = lane; e < c * k; e += kWave) v.Lp[m[FM_L] + e] = Mg[e];
        for (int i = lane; i < c; i += kWave) v.fv[m[FM_F] + i] = i < k ? SA.x[first + i] : T(0);
    }
    __syncthreads();
    // update vectors of children outside the task
    for (int lf = w; lf < v.nf; lf += nw) {
        const int* m = v.mt + lf * kFM;
        for (int q = m[FM_CH]; q < m[FM_CH] + m[FM_NCH]; ++q) {
            const int cv = v.chl[q];
            if (cv >= 0) continue;
            const int ch = -cv - 1;
            const int u = SA.sn_c[ch] - SA.sn_k[ch];
            const T* uc = SA.ubuf + SA.uoff[ch];
            const int32_t* rmg = SA.rmap + SA.rmoff[ch];
            for (int a = lane; a < u; a += kWave) v.fv[m[FM_F] + rmg[a]] += uc[a];
            wave_sync();
        }
    }
    __syncthreads();
    if (FLAGS) {
        for (int lf = w; lf < v.nf; lf += nw) {  // fronts are listed children first
            const int* m = v.mt + lf * kFM;
            for (int q = m[FM_CH]; q < m[FM_CH] + m[FM_NCH]; ++q)
                if (v.chl[q] >= 0) lds_wait(v.flg + v.chl[q], A.st);
            fwd_lds_front<T, FLAGS>(v, lf, lane);
            lds_publish(v.flg + lf, lane);
        }
        __syncthreads();
    } else {
        for (int L = 0; L < v.nlev; ++L) {
            for (int lf = v.lv[L] + w; lf < v.lv[L + 1]; lf += nw) fwd_lds_front<T, FLAGS>(v, lf, lane);
            __syncthreads();
        }
    }
    for (int lf = w; lf < v.nf; lf += nw) {
        const int* m = v.mt + lf * kFM;
        const int s = m[FM_S], c = m[FM_C], k = m[FM_K], first = m[FM_FIRST];
        const T* f = v.fv + m[FM_F];
        for (int j = lane; j < k; j += kWave) SA.x[first + j] = f[j];
        if (m[FM_EXTPAR]) {
            T* us = SA.ubuf + SA.uoff[s];
            for (int a = lane; a < c - k; a += kWave) us[a] = f[k + a];
        }
    }
}

template <typename T>
__device__ inline void bwd_lds_front(const LdsView<T>& v, int lf, int lane) {
    const int* m = v.mt + lf * kFM;
    const int c = m[FM_C], k = m[FM_K];
    T* xv = v.fv + m[FM_F];
    const T* Lp = v.Lp + m[FM_L];
    const int* sr = v.srl + m[FM_SRC];
    for (int i = k + lane; i < c; i += kWave) {
        const int code = sr[i - k];
        if (code >= 0) xv[i] = v.fv[code];  // ancestor inside the task, solved before
    }
    wave_sync();
    // x_s = M^T [y_s; x_rows]  (lane j < k: one column of M)
    double out = 0.0;
    if (lane < k)
        for (int i = lane; i < c; ++i) out += (double)Lp[lane * c + i] * (double)xv[i];
    wave_sync();
    if (lane < k) xv[lane] = (T)out;
    wave_sync();
}

template <typename T, bool FLAGS>
__global__ __launch_bounds__(1024) void k_bwd_lds(LdsTaskArgs A, SolveArgs<T> SA, int task_base) {
    extern __shared__ unsigned char smem_raw[];
    const int32_t* H = A.th + (size_t)(task_base + blockIdx.x) * kTH;
    LdsView<T> v = lds_task_load<T>(A, H, smem_raw, true);
    const int nw = blockDim.x / kWave, w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
    __syncthreads();
    for (int lf = w; lf < v.nf; lf += nw) {
        const int* m = v.mt + lf * kFM;
        const int s = m[FM_S], c = m[FM_C], k = m[FM_K], first = m[FM_FIRST];
        const T* Mg = SA.M + SA.moff[s];
        for (int e = lane; e < c * k; e += kWave) v.Lp[m[FM_L] + e] = Mg[e];
        T* xv = v.fv + m[FM_F];
        const int* sr = v.srl + m[FM_SRC];
        for (int i = lane; i < c; i += kWave) {
            if (i < k) {
                xv[i] = SA.x[first + i];
            } else {
                const int code = sr[i - k];
                if (code < 0) xv[i] = SA.x[-code - 1];  // ancestor outside the task: final
            }
        }
    }
    __syncthreads();
    if (FLAGS) {
        for (int lf = v.nf - 1 - w; lf >= 0; lf -= nw) {  // parents before children
            const int par = v.mt[lf * kFM + FM_PAR];
            if (par >= 0) lds_wait(v.flg + par, A.st);
            bwd_lds_front<T>(v, lf, lane);
            lds_publish(v.flg + lf, lane);
        }
        __syncthreads();
    } else {
        for (int L = v.nlev - 1; L >= 0; --L) {
            for (int lf = v.lv[L] + w; lf < v.lv[L + 1]; lf += nw) bwd_lds_front<T>(v, lf, lane);
            __syncthreads();
        }
    }
    for (int lf = w; lf < v.nf; lf += nw) {
        const int* m = v.mt + lf * kFM;
        const int k = m[FM_K], first = m[FM_FIRST];
        const T* xv = v.fv + m[FM_F];
        for (int j = lane; j < k; j += kWave) SA.x[first + j] = xv[j];
    }
}

// ---------------------------------------------------------------------------
// Partitioned-inverse top solve (the "partitioned sparse inverse" of Alvarado,
// Yu and Betancourt, IEEE Trans. Power Systems 5(2), 1990).  The supernodal
// tree is split into bottom subtrees (solved by the LDS tasks above) and the
// top T, which is closed under taking parents.  Because no bottom column is an
// ancestor of a top column, L = [L_BB 0; L_TB L_TT] and
//   forward:  c_T = b_T + (update vectors of the bottom roots),  y_T = L_TT^-1 c_T
//   backward: x_T = L_TT^-T y_T,  then the bottom tasks finish x_B.
// L_TT^-1 is stored explicitly: column j is non-zero only on the ancestors of j
// (its path to the root), so each top supernode s owns a dense block of
// P_s x k_s values (P_s = columns on its path), recomputed after every
// factorization by walking the path once (k_pi_inv).  The two triangular
// sweeps of the top -- dozens of dependent levels, each a launch or a barrier
// -- become two independent gathers.
// ---------------------------------------------------------------------------
struct PiArgs {
    const int32_t* sn;      // top supernodes in a postorder of the top forest
    const int32_t* plen;    // P_s, path rows of each top supernode
    const int64_t* boff;    // offset of each supernode's block
    const int64_t* poff;    // offset of each supernode's path rows in ploc
    const int32_t* ploc;    // top-local index of each path row
    const int32_t* parent;  // supernode parent (-1 at a root)
    const int32_t* tcol;    // column of each top-local index
    const int32_t* colsn;   // top supernode (local) of each top-local column
    int ntop;               // top supernodes
    int ntc;                // top columns
    int slice;              // LDS elements of one per-wave buffer (cmax * kmax)
    // per top column: start of its block column, of its path rows, path length, column in its supernode
    const int64_t* cboff;
    const int64_t* cpoff;
    const int32_t* cplen;
    const int32_t* cjj;
    // per top supernode: its first column and its column count
    const int32_t* ufirst;
    const int32_t* uk;
    // units of the backward product (supernode, first column), longest path first
    const int32_t* bunit;
    int nunit;
    // per top supernode: the fronts on its path to the root, upath[upoff[u] .. upoff[u+1])
    const int32_t* upath;
    const int32_t* upoff;
};

// One wave per top supernode s: the k_s unit right-hand sides of its columns are
// carried up the path front by front.  At front p the block F (c_p x k_s) becomes
// M_p F_top + [0; F_bot] with the front's solve matrix M_p = [L11^-1; -L21 L11^-1]:
// its first k_p rows are the block rows of L_TT^-1 for p's columns, the rest is
// extend-added (through rmap) into the parent's block.  The path's fronts come
// from the plan: their sizes and offsets are loaded lane-parallel (a lane per
// front), and the next front's M_p is loaded into registers during the
// current front's product, so a step waits for LDS, not for memory.
constexpr int kPiInvPre = 8;  // M_p entries per lane prefetched (fronts of c k <= 512)

template <typename T>
__global__ __launch_bounds__(256) void k_pi_inv(PiArgs A, SolveArgs<T> SA, T* __restrict__ B) {
    extern __shared__ unsigned char smem_raw[];
    const int nw = blockDim.x / kWave, w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
    const int u = blockIdx.x * nw + w;
    if (u >= A.ntop) return;  // the whole wave leaves; no workgroup barrier below
    T* F0 = reinterpret_cast<T*>(smem_raw) + (size_t)w * 3 * A.slice;
    T* F1 = F0 + A.slice;
    T* Mp = F1 + A.slice;
    const int ks = A.uk[u];
    const int Ps = A.plen[u];
    T* Bs = B + A.boff[u];
    const int32_t* path = A.upath + A.upoff[u];
    const int nf = A.upoff[u + 1] - A.upoff[u];
    // sizes and offsets of the path's fronts [chunk, chunk + 64), lane f for front chunk + f
    int chunk = 0, c_l = 0, k_l = 0;
    int64_t mo_l = 0, ro_l = 0;
    auto fetch = [&](int base) {
        chunk = base;
        const int f = base + lane;
        if (f < nf) {
            const int p = path[f];
            c_l = SA.sn_c[p];
            k_l = SA.sn_k[p];
            mo_l = SA.moff[p];
            ro_l = SA.rmoff[p];
        }
    };
    auto desc = [&](int f, int& c, int& k, int64_t& mo, int64_t& ro) {
        if (f - chunk >= kWave) fetch(f);  // uniform: the next 64 fronts of a long path
        c = __shfl(c_l, f - chunk, kWave);
        k = __shfl(k_l, f - chunk, kWave);
        mo = __shfl(mo_l, f - chunk, kWave);
        ro = __shfl(ro_l, f - chunk, kWave);
    };
    fetch(0);
    int c, k;
    int64_t mo, ro;
    desc(0, c, k, mo, ro);
    T mv[kPiInvPre];
    auto prefetch = [&](int64_t moff, int n) {
#pragma unroll
        for (int q = 0; q < kPiInvPre; ++q) {
            const int e = lane + q * kWave;
            mv[q] = e < n ? SA.M[moff + e] : T(0);
        }
    };
    prefetch(mo, c * k);
    // F is c x ks, row-major: F[i * ks + jj]
    for (int e = lane; e < c * ks; e += kWave) {
        const int i = e / ks, jj = e - i * ks;
        F0[e] = i == jj ? T(1) : T(0);
    }
    int ro_b = 0;  // block rows written so far
    for (int f = 0;; ++f) {
        const bool last = f + 1 >= nf;
        int cn = 0, kn = 0;
        int64_t mon = 0, ron = 0;
        if (!last) desc(f + 1, cn, kn, mon, ron);
        // this front's M_p (c x k, column-major) into LDS: the prefetched part, then any rest
#pragma unroll
        for (int q = 0; q < kPiInvPre; ++q) {
            const int e = lane + q * kWave;
            if (e < c * k) Mp[e] = mv[q];
        }
        for (int e = lane + kPiInvPre * kWave; e < c * k; e += kWave) Mp[e] = SA.M[mo + e];
        for (int e = lane; e < cn * ks; e += kWave) F1[e] = T(0);
        if (!last) prefetch(mon, cn * kn);  // the next front's, in flight during this product
        wave_sync();
        const int32_t* rm = SA.rmap + ro;
        for (int e = lane; e < c * ks; e += kWave) {
            const int i = e / ks, jj = e - i * ks;
            T acc = i < k ? T(0) : F0[e];
            for (int j = 0; j < k; ++j) acc += Mp[j * c + i] * F0[j * ks + jj];
            if (i < k)
                Bs[(ro_b + i) + (int64_t)Ps * jj] = acc;
            else
                F1[rm[i - k] * ks + jj] = acc;
        }
        ro_b += k;
        if (last) break;
        wave_sync();
        T* t = F0;
        F0 = F1;
        F1 = t;
        c = cn;
        k = kn;
        mo = mon;
        ro = ron;
    }
}

// Row-grouped copy of the blocks for the forward product.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_pi_gather(int64_t nq, const int64_t* __restrict__ pos,
                                                     const T* __restrict__ B, T* __restrict__ val) {
    const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (q < nq) val[q] = B[pos[q]];
}

// c_T = b_T + update vectors of the bottom subtree roots (fixed order per row).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_pi_assemble(int ntc, const int32_t* __restrict__ tcol,
                                                       const int32_t* __restrict__ cptr,
                                                       const int64_t* __restrict__ cpos, const T* __restrict__ x,
                                                       const T* __restrict__ ubuf, T* __restrict__ ctop) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= ntc) return;
    T acc = x[tcol[t]];
    for (int q = cptr[t]; q < cptr[t + 1]; ++q) acc += ubuf[cpos[q]];
    ctop[t] = acc;
}

// entries per segment of the forward product (a multiple of 64): 256 keeps
// enough waves in flight (1024 measured 35% slower at 200k buses)
constexpr int kPiSegMax = 256;
constexpr int kPiSegWave = 4;  // forward-product segments per wave

// Forward product y_T = L_TT^-1 c_T, kPiSegWave segments of a row's entries per
// wave: their headers are wave-uniform (scalar loads), then every lane issues
// the loads of all of them before the first product, so a wave waits for
// memory twice.  The entries of a row are one contiguous range of c_T (the top
// fronts are numbered in postorder), so entry q multiplies c_T[segc + q]: the
// matrix values and c_T stream side by side, with no column index per entry.
// Out of range, a lane loads the segment's first entry and adds 0.  A row held
// by one segment (most rows) is written directly (segdst >= 0); longer rows
// leave partial sums for k_pi_rowsum (segdst = ~segment).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_pi_fwd(int nseg, const int64_t* __restrict__ segq,
                                                  const int64_t* __restrict__ segc, const int32_t* __restrict__ segdst,
                                                  const T* __restrict__ val, const T* __restrict__ ctop,
                                                  double* __restrict__ part, T* __restrict__ ytop) {
    const int g0 =
        __builtin_amdgcn_readfirstlane((blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave) * kPiSegWave);
    const int lane = threadIdx.x & (kWave - 1);
    if (g0 >= nseg) return;
    constexpr int kU = kPiSegMax / kWave;
    int64_t q0[kPiSegWave], q1[kPiSegWave];
    const T* cq[kPiSegWave];
#pragma unroll
    for (int u = 0; u < kPiSegWave; ++u) {
        const int g = min(g0 + u, nseg - 1);
        q0[u] = segq[g];
        q1[u] = g0 + u < nseg ? segq[g + 1] : q0[u];
        cq[u] = ctop + segc[g];  // c_T aligned with the entries
    }
    T v[kPiSegWave][kU], c[kPiSegWave][kU];
#pragma unroll
    for (int u = 0; u < kPiSegWave; ++u) {
#pragma unroll
        for (int k = 0; k < kU; ++k) {
            const int64_t q = q0[u] + lane + (int64_t)k * kWave;
            const int64_t qq = q < q1[u] ? q : q0[u];
            v[u][k] = val[qq];
            c[u][k] = cq[u][qq];
        }
    }
#pragma unroll
    for (int u = 0; u < kPiSegWave; ++u) {
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < kU; ++k)
            acc += q0[u] + lane + (int64_t)k * kWave < q1[u] ? (double)v[u][k] * (double)c[u][k] : 0.0;
        acc = wave_sum(acc);
        if (lane == 0 && g0 + u < nseg) {
            const int d = segdst[g0 + u];
            if (d >= 0)
                ytop[d] = (T)acc;
            else
                part[~d] = acc;
        }
    }
}

// Rows spread over several segments: one wave per such row.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_pi_rowsum(int nrows, const int32_t* __restrict__ rows,
                                                     const int32_t* __restrict__ segptr,
                                                     const double* __restrict__ part, T* __restrict__ ytop) {
    const int q = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    if (q >= nrows) return;  // uniform per wave
    const int t = rows[q];
    double acc = 0.0;
    for (int g = segptr[t] + lane; g < segptr[t + 1]; g += kWave) acc += part[g];
    acc = wave_sum(acc);
    if (lane == 0) ytop[t] = (T)acc;
}

// Backward product x_T = L_TT^-T y_T for FP64 solve data: one wave per top
// column, a dot product of its (contiguous) block column with y_T gathered
// along the path (measured faster than the units below for FP64 blocks: 16.7
// against 20.1 us at 200k buses; the units hold 106 VGPRs in FP64).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_pi_bwd_col(PiArgs A, SolveArgs<T> SA, const T* __restrict__ B,
                                                  const T* __restrict__ ytop) {
    const int t = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    if (t >= A.ntc) return;  // uniform per wave
    // per-column metadata, one round trip: block column, path rows, path length
    const int col = A.tcol[t], jj = A.cjj[t], Ps = A.cplen[t];
    const T* Bc = B + A.cboff[t];
    const int32_t* pl = A.ploc + A.cpoff[t];
    double acc = 0.0;
    constexpr int kU = 4;  // loads of a lane in flight together
    for (int r0 = jj + lane; r0 < Ps; r0 += kU * kWave) {
        double bv[kU];
        int c[kU];
#pragma unroll
        for (int k = 0; k < kU; ++k) {
            const int r = r0 + k * kWave;
            bv[k] = r < Ps ? (double)Bc[r] : 0.0;
            c[k] = r < Ps ? pl[r] : -1;
        }
#pragma unroll
        for (int k = 0; k < kU; ++k) acc += c[k] >= 0 ? bv[k] * (double)ytop[c[k]] : 0.0;
    }
    acc = wave_sum(acc);
    if (lane == 0) SA.x[col] = (T)acc;
}

// Backward product x_T = L_TT^-T y_T for FP32 solve data: one wave per unit of
// a top supernode's columns (kPiBwdCols of them, units longest path first;
// 14.8 -> 13.7 us against one wave per column at 200k buses).  The columns share
// the path rows, so y_T is gathered along the path once for the unit, each
// column with its own partial sum (column jj of the block starts at row jj;
// rows above are zero).  Out of range, a lane loads row j0 of column j0 and
// adds 0.
constexpr int kPiBwdRows = 8;  // path rows of a lane in flight together
constexpr int kPiBwdCols = 4;  // columns per unit
constexpr int kPiUnitShift = 7;  // unit = supernode << kPiUnitShift | first column of the unit

template <typename T>
__global__ __launch_bounds__(kBlock) void k_pi_bwd(PiArgs A, SolveArgs<T> SA, const T* __restrict__ B,
                                                  const T* __restrict__ ytop) {
    const int w = __builtin_amdgcn_readfirstlane(blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave);
    const int lane = threadIdx.x & (kWave - 1);
    if (w >= A.nunit) return;  // uniform per wave
    const int unit = A.bunit[w];
    const int u = unit >> kPiUnitShift, j0 = unit & ((1 << kPiUnitShift) - 1);
    // per-supernode metadata, one round trip: first column, columns, path rows, block, path
    const int col0 = A.ufirst[u], ks = A.uk[u], Ps = A.plen[u];
    const T* Bs = B + A.boff[u];
    const int32_t* pl = A.ploc + A.poff[u];
    double acc[kPiBwdCols];
#pragma unroll
    for (int jj = 0; jj < kPiBwdCols; ++jj) acc[jj] = 0.0;
    for (int r0 = j0 + lane; r0 < Ps; r0 += kPiBwdRows * kWave) {
        int c[kPiBwdRows];
        T bv[kPiBwdCols][kPiBwdRows];
#pragma unroll
        for (int k = 0; k < kPiBwdRows; ++k) {
            const int r = r0 + k * kWave;
            c[k] = pl[r < Ps ? r : j0];
        }
#pragma unroll
        for (int jj = 0; jj < kPiBwdCols; ++jj) {
            const int64_t cj = (int64_t)Ps * min(j0 + jj, ks - 1);
#pragma unroll
            for (int k = 0; k < kPiBwdRows; ++k) {
                const int r = r0 + k * kWave;
                bv[jj][k] = Bs[cj + (r < Ps ? r : j0)];
            }
        }
        T y[kPiBwdRows];
#pragma unroll
        for (int k = 0; k < kPiBwdRows; ++k) y[k] = ytop[c[k]];
#pragma unroll
        for (int jj = 0; jj < kPiBwdCols; ++jj) {
#pragma unroll
            for (int k = 0; k < kPiBwdRows; ++k) {
                const int r = r0 + k * kWave;
                acc[jj] += (r < Ps && r >= j0 + jj && j0 + jj < ks) ? (double)bv[jj][k] * (double)y[k] : 0.0;
            }
        }
    }
#pragma unroll
    for (int jj = 0; jj < kPiBwdCols; ++jj) {
        if (j0 + jj < ks) {  // uniform per wave
            const double v = wave_sum(acc[jj]);
            if (lane == 0) SA.x[col0 + j0 + jj] = (T)v;
        }
    }
}

// ---------------------------------------------------------------------------
// Dense solve matrices of the bottom subtrees.  A bottom subtree t with n
// columns (its fronts' columns, children first) and u update rows (the rows of
// its root beyond the root's columns; all of them lie in the top) has
//   M_t = [L_BB^-1 ; -L_UB L_BB^-1]      (R = n + u rows, n columns)
// so that its forward step is one product [y_B ; u_t] = M_t b_B (y_B the
// solution of its columns, u_t the update vector its root hands to the top)
// and its backward step is x_B = M_t^T [y_B ; x_U].  The dependent chain of
// front steps inside the subtree becomes one independent product per subtree.
// M_t is stored column by column from the diagonal down (lower trapezoid:
// column j holds rows j..R-1), n <= 64, R <= 128.
// ---------------------------------------------------------------------------
constexpr int kSubCols = 64, kSubRows = 128;
static_assert(kSubRows <= 2 * kWave, "the backward gather takes rows lane and lane + 64");
constexpr int kSubBins = 4;  // launch groups of the subtree-matrix kernels

struct SubArgs {
    int nsub;
    int64_t msub;         // entries of all M_t (the row-major copies follow)
    const int32_t* root;  // root supernode
    const int32_t* n;     // columns
    const int32_t* R;     // columns + update rows
    // L of subtree t's fronts as rows of structural entries (CSR over the dense
    // rows, columns ascending): row r holds the dense columns set in
    // gmask[gmo[t] + r], from grs[gmo[t] + t + r].  The gather's entries e in
    // [geptr[t], geptr[t+1]), column-major within each front (neighbouring
    // loads touch neighbouring addresses): gsrc[e] their position from
    // gbase[t] in the front storage, gdst[e] their position in the CSR.
    const int64_t* geptr;
    const int64_t* gbase;
    const int32_t* gsrc;
    const int32_t* gdst;
    const int64_t* gmo;
    const uint64_t* gmask;
    const int32_t* grs;
    const int32_t* cptr;   // columns of subtree t: cols[cptr[t] .. cptr[t] + n[t])
    const int32_t* cols;
    const int64_t* moff;   // start of M_t
    const int32_t* acptr;  // per subtree column: the update positions that add into its right-hand side
    const int64_t* acpos;
    const int64_t* uoff;  // per subtree: its root's update rows in the update buffer
    const int64_t* urow;  // and their row indices in the front row lists
};

__host__ __device__ inline int64_t sub_col_off(int j, int R) { return (int64_t)j * R - (int64_t)j * (j - 1) / 2; }

// Row-major copy of M_t (row i holds columns 0..min(i, n-1)) for the backward
// product; it sits msub entries after the column-major one.
__host__ __device__ inline int64_t sub_row_off(int i, int n) {
    return i <= n ? (int64_t)i * (i + 1) / 2 : (int64_t)n * (n + 1) / 2 + (int64_t)(i - n) * n;
}

// Load at a 32-bit element offset from a wave-uniform base: the address stays a
// scalar base plus one 32-bit register instead of a 64-bit pair per load.
template <typename T>
__device__ inline T ld_off(const T* base, uint32_t i) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + (size_t)(i * (uint32_t)sizeof(T)));
}

// After each factorization: M_t with all columns at once, the subtree's L
// gathered into LDS.  Lane j carries column j through the
// row-oriented substitution
//   X[i][j] = (delta_ij - sum_{k<i} L[i][k] X[k][j]) / L[i][i],   i < n,
//   M[n + r][j] = -sum_{k<n} L[n + r][k] X[k][j],
// whose inner sums are independent LDS loads (broadcast L, conflict-free X)
// rather than a chain of cross-lane steps.  The sums run over the structural
// entries of row i only: L sits in LDS as rows of its entries with a 64-bit
// column mask per row (from the plan), X packed lower triangular, so a
// subtree needs a fraction of the LDS of dense L and X and a CU holds several
// (200k buses: 609 -> 294 us per refactor; most rows cost one round of LDS
// loads instead of i / 8).  The rows of X depend on each other: one wave
// walks them; the update rows n + r do not: once X is there, the block's
// waves share them.
// Row i of L times X over row i's structural entries (columns: the set bits
// of m, wave-uniform; values: consecutive from Lr): eight at a time, eight
// independent partial sums, so a row with at most eight entries before the
// diagonal (most rows) waits for LDS once.  X is packed lower triangular (row
// k holds columns 0..k); a slot past the last bit, or a lane past row k,
// reads an entry of row 0 (computed, finite) and adds 0.
__device__ inline int xoff(int k) { return k * (k + 1) / 2; }

__device__ inline double sub_m_row(const double* Lr, const double* X, int j, uint64_t m) {
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int q = 0; m; q += 8) {
        int k[8];
        bool ok[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            ok[u] = m != 0;
            k[u] = ok[u] ? __builtin_ctzll(m) : 0;
            m &= m - 1;
        }
        double lv[8], xv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const bool in = ok[u] && j <= k[u];
            lv[u] = Lr[ok[u] ? q + u : 0];
            xv[u] = X[in ? xoff(k[u]) + j : 0];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) acc[u] -= (ok[u] && j <= k[u]) ? lv[u] * xv[u] : 0.0;
    }
    return ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
}

__device__ inline uint64_t uniform_u64(uint64_t v) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return (uint64_t)hi << 32 | lo;
}

// LDS of one subtree: its L entries and packed X (FP64), row masks, row starts
__host__ __device__ inline size_t sub_m_lds(int64_t nnz, int n, int R) {
    return (size_t)(nnz + (int64_t)n * (n + 1) / 2 + R) * sizeof(double) + (size_t)(R + 2) * sizeof(int32_t);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_sub_m_rows(SubArgs B, SolveArgs<T> A, const int32_t* __restrict__ list,
                                                      T* __restrict__ Mout) {
    extern __shared__ double sm[];
    const int t = list[blockIdx.x];
    const int n = B.n[t], R = B.R[t];
    const int nw = kBlock / kWave, w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
    const int64_t e0 = B.geptr[t], e1 = B.geptr[t + 1];  // e1 > e0: the diagonal at least
    const int nnz = (int)(e1 - e0);
    double* Lc = sm;                                               // the entries, row by row
    double* X = Lc + nnz;                                          // packed lower triangular
    uint64_t* mk = reinterpret_cast<uint64_t*>(X + n * (n + 1) / 2);  // per row, its columns
    int32_t* rs = reinterpret_cast<int32_t*>(mk + R);              // per row, its first entry
    const int64_t mo = B.gmo[t];
    for (int r = threadIdx.x; r <= R; r += kBlock) {
        if (r < R) mk[r] = B.gmask[mo + r];
        rs[r] = B.grs[mo + t + r];
    }
    // the gather: every thread issues a batch of entry loads before their LDS stores
    constexpr int GB = 8;
    const T* Lb = A.F + B.gbase[t];
    for (int64_t e = e0 + threadIdx.x; e < e1; e += (int64_t)GB * kBlock) {
        int sv[GB], dv[GB];
#pragma unroll
        for (int u = 0; u < GB; ++u) {
            const int64_t ee = e + (int64_t)u * kBlock;
            const int64_t ec = ee < e1 ? ee : e0;
            sv[u] = B.gsrc[ec];
            dv[u] = B.gdst[ec];
        }
        T v[GB];
#pragma unroll
        for (int u = 0; u < GB; ++u) v[u] = Lb[sv[u]];
#pragma unroll
        for (int u = 0; u < GB; ++u)
            if (e + (int64_t)u * kBlock < e1) Lc[dv[u]] = (double)v[u];
    }
    __syncthreads();
    T* Mt = Mout + B.moff[t];
    T* Mr = Mout + B.msub + B.moff[t];
    const int j = lane;
    if (w == 0) {  // the rows of X, in order
        for (int i = 0; i < n; ++i) {
            const int r0 = __builtin_amdgcn_readfirstlane(rs[i]), r1 = __builtin_amdgcn_readfirstlane(rs[i + 1]);
            const uint64_t m = uniform_u64(mk[i]) & ((1ull << i) - 1);  // entries before the diagonal
            double v = sub_m_row(Lc + r0, X, j, m);
            if (i == j) v += 1.0;
            v = v / Lc[r1 - 1];  // the diagonal: the row's last entry
            if (j <= i) {
                X[xoff(i) + j] = v;
                Mt[sub_col_off(j, R) + i - j] = (T)v;
                Mr[sub_row_off(i, n) + j] = (T)v;
            }
        }
    }
    __syncthreads();
    for (int i = n + w; i < R; i += nw) {  // the update rows, shared by the waves
        const int r0 = __builtin_amdgcn_readfirstlane(rs[i]);
        const double v = sub_m_row(Lc + r0, X, j, uniform_u64(mk[i]));
        if (j < n) {
            Mt[sub_col_off(j, R) + i - j] = (T)v;
            Mr[sub_row_off(i, n) + j] = (T)v;
        }
    }
}

constexpr int kSubBatch = 8;    // column loads in flight per lane and row (16 measured slower)
constexpr int kSubBatchB = 16;  // row loads in flight per lane, backward

// Forward: one wave per subtree; lane i accumulates rows i and i + 64 over the
// columns of M_t, kSubBatch coalesced column loads at a time (2 kSubBatch when
// every row fits one lane).
// LEVEL 0: bottom subtrees, 1: middle subtrees (one instantiation each, so
// that profiles tell the two launches apart)
template <typename T, int LEVEL>
__device__ __forceinline__ void sub_fwd_wave(const SubArgs& B, const SolveArgs<T>& A, const T* __restrict__ M,
                                             const int t, const int lane, double* bs) {
    const int n = B.n[t], R = B.R[t];
    const int32_t* cols = B.cols + B.cptr[t];
    const T* Mt = M + B.moff[t];
    T* u = A.ubuf + B.uoff[t];
    // entry (i, j) of the packed column-major M_t: j R - j (j - 1) / 2 + i - j
    auto cidx = [&](int j, int i) { return (uint32_t)(j * R - j * (j - 1) / 2 + i - j); };
    // FP32 elements: unconditional loads (an entry outside the trapezoid reads
    // entry 0) of raw values, selected and converted in the product; a load
    // under a branch gets its conversion in the same branch and waits for it
    // there, which serialized the batch.  The column index goes first, so the
    // gather waits for it alone (loads complete in issue order).  FP64
    // elements keep predicated loads and the gather after the first batch
    // (measured faster for them).
    constexpr bool clamp = sizeof(T) == 4;
    int cl = clamp && lane < n ? cols[lane] : 0;
    // the right-hand side of the subtree columns into LDS (issued after the
    // first column loads, which need only the header)
    auto gather = [&]() {
        if (lane < n) {
            double b = (double)A.x[cl];
            if (LEVEL == 1) {  // a middle subtree: the bottom roots' update rows add in first
                const int c = B.cptr[t] + lane;
                for (int q = B.acptr[c]; q < B.acptr[c + 1]; ++q) b += (double)A.ubuf[B.acpos[q]];
            }
            bs[lane] = b;
        }
        wave_sync();
    };
    double a0 = 0.0, a1 = 0.0;
    if (R <= kWave) {
        // one row per lane: twice the columns per batch in the same registers
        constexpr int NB = 2 * kSubBatch;
        T m[NB];
        auto ok = [&](int j, int i) { return j < n && i >= j && i < R; };
        auto load = [&](int j0) {
#pragma unroll
            for (int q = 0; q < NB; ++q)
                m[q] = clamp ? ld_off(Mt, ok(j0 + q, lane) ? cidx(j0 + q, lane) : 0u)
                             : (ok(j0 + q, lane) ? ld_off(Mt, cidx(j0 + q, lane)) : T(0));
        };
        load(0);
        if (!clamp && lane < n) cl = cols[lane];
        gather();
        for (int j0 = 0; j0 < n; j0 += NB) {
            if (j0) load(j0);
#pragma unroll
            for (int q = 0; q < NB; ++q)
                a0 += (!clamp || ok(j0 + q, lane) ? (double)m[q] : 0.0) * (j0 + q < n ? bs[j0 + q] : 0.0);
        }
    } else {
        // rows lane and lane + 64
        T m0[kSubBatch], m1[kSubBatch];
        auto ok = [&](int j, int i) { return j < n && i >= j && i < R; };
        auto load = [&](int j0) {
#pragma unroll
            for (int q = 0; q < kSubBatch; ++q) {
                const int j = j0 + q;
                m0[q] = clamp ? ld_off(Mt, ok(j, lane) ? cidx(j, lane) : 0u)
                              : (ok(j, lane) ? ld_off(Mt, cidx(j, lane)) : T(0));
                m1[q] = clamp ? ld_off(Mt, ok(j, lane + kWave) ? cidx(j, lane + kWave) : 0u)
                              : (ok(j, lane + kWave) ? ld_off(Mt, cidx(j, lane + kWave)) : T(0));
            }
        };
        load(0);
        if (!clamp && lane < n) cl = cols[lane];
        gather();
        for (int j0 = 0; j0 < n; j0 += kSubBatch) {
            if (j0) load(j0);
#pragma unroll
            for (int q = 0; q < kSubBatch; ++q) {
                const double b = j0 + q < n ? bs[j0 + q] : 0.0;
                a0 += (!clamp || ok(j0 + q, lane) ? (double)m0[q] : 0.0) * b;
                a1 += (!clamp || ok(j0 + q, lane + kWave) ? (double)m1[q] : 0.0) * b;
            }
        }
    }
    if (lane < n)
        A.x[cl] = (T)a0;
    else if (lane < R)
        u[lane - n] = (T)a0;
    if (lane + kWave < R) u[lane + kWave - n] = (T)a1;
}

template <typename T, int LEVEL>
__global__ __launch_bounds__(kBlock) void k_sub_fwd(SubArgs B, SolveArgs<T> A, const T* __restrict__ M, int t0,
                                                   int t1) {
    __shared__ double bsh[kBlock / kWave][kSubCols];
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int t = t0 + blockIdx.x * (kBlock / kWave) + w;
    const int lane = threadIdx.x & (kWave - 1);
    if (t >= t1) return;  // uniform per wave
    sub_fwd_wave<T, LEVEL>(B, A, M, t, lane, bsh[w]);
}

// Backward: one wave per subtree; lane j accumulates column j over the rows of
// the row-major copy, kSubBatchB coalesced row loads at a time.
template <typename T, int LEVEL>
__device__ __forceinline__ void sub_bwd_wave(const SubArgs& B, const SolveArgs<T>& A, const T* __restrict__ Mrow,
                                             const int t, const int lane, double* v) {
    const int n = B.n[t], R = B.R[t];
    const int32_t* cols = B.cols + B.cptr[t];
    const int32_t* urow = A.rows + B.urow[t];
    const T* Mt = Mrow + B.moff[t];
    T mv[kSubBatchB];  // raw values; FP32 loads unconditional (see k_sub_fwd)
    constexpr bool clamp = sizeof(T) == 4;
    auto load = [&](int i0) {  // row loads of batch i0 (the first one overlaps the gather below)
#pragma unroll
        for (int u = 0; u < kSubBatchB; ++u) {
            const int i = i0 + u;
            const uint32_t ro = i <= n ? i * (i + 1) / 2 : n * (n + 1) / 2 + (i - n) * n;  // sub_row_off
            const bool ok = i < R && lane <= i && lane < n;
            mv[u] = clamp ? ld_off(Mt, ok ? ro + lane : 0u) : (ok ? ld_off(Mt, ro + lane) : T(0));
        }
    };
    // FP32: the gather's indices first (loads complete in issue order), then
    // the first row batch, then the gather itself (R <= 2 * 64); FP64: the
    // first row batch, then the gather
    if (clamp) {
        const int ia = lane, ib = lane + kWave;
        const int xa = ia < R ? (ia < n ? cols[ia] : urow[ia - n]) : 0;
        const int xb = ib < R ? (ib < n ? cols[ib] : urow[ib - n]) : 0;
        load(0);
        if (ia < R) v[ia] = (double)A.x[xa];
        if (ib < R) v[ib] = (double)A.x[xb];
    } else {
        load(0);
        for (int i = lane; i < R; i += kWave) v[i] = (double)(i < n ? A.x[cols[i]] : A.x[urow[i - n]]);
    }
    wave_sync();
    double acc = 0.0;
    for (int i0 = 0; i0 < R; i0 += kSubBatchB) {
        if (i0) load(i0);
#pragma unroll
        for (int u = 0; u < kSubBatchB; ++u) {
            const int i = i0 + u;
            acc += (!clamp || (i < R && lane <= i && lane < n) ? (double)mv[u] : 0.0) * (i < R ? v[i] : 0.0);
        }
    }
    if (lane < n) A.x[cols[lane]] = (T)acc;
}

template <typename T, int LEVEL>
__global__ __launch_bounds__(kBlock) void k_sub_bwd(SubArgs B, SolveArgs<T> A, const T* __restrict__ Mrow, int t0,
                                                   int t1) {
    __shared__ double vsh[kBlock / kWave][kSubRows];
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int t = t0 + blockIdx.x * (kBlock / kWave) + w;
    const int lane = threadIdx.x & (kWave - 1);
    if (t >= t1) return;  // uniform per wave
    sub_bwd_wave<T, LEVEL>(B, A, Mrow, t, lane, vsh[w]);
}

// Bottom-subtree products with the small subtrees packed: a schedule lists
// the waves (largest first), each either one subtree (more than 32 rows: the
// one-wave code above) or two subtrees of at most 32 rows, or four of at most
// 16, side by side on G = 32 or 16 lanes.  58 % of the bottom subtrees of the
// 200k-bus grid have at most 32 rows: packing them cuts the waves of a product
// by a third (13,178 -> 8,400).
struct SubPack {
    int nw;               // waves
    const int32_t* kind;  // per wave: 64 (one subtree), 32 or 16 (lanes per subtree)
    const int32_t* sub;   // per wave: four subtree indices (-1: none)
};

constexpr int kPackBatch = 16;  // column (forward) or row (backward) loads in flight per lane

// Forward product of subtree t (t < 0: none) on G lanes, lane l = row l
// (R <= G); bs: the G doubles of LDS of this lane group.
template <typename T, int G>
__device__ __forceinline__ void sub_fwd_packed(const SubArgs& B, const SolveArgs<T>& A, const T* __restrict__ M,
                                               const int t, const int l, double* bs) {
    constexpr bool clamp = sizeof(T) == 4;  // unconditional FP32 loads, as in sub_fwd_wave
    int n = 0, R = 0, cp = 0;
    int64_t mo = 0, uo = 0;
    if (t >= 0) {
        n = B.n[t];
        R = B.R[t];
        cp = B.cptr[t];
        mo = B.moff[t];
        uo = B.uoff[t];
    }
    const T* Mt = M + mo;
    const int cl = l < n ? B.cols[cp + l] : 0;  // issued first: the gather waits for it alone
    auto ok = [&](int j) { return j < n && l >= j && l < R; };
    auto cidx = [&](int j) { return (uint32_t)(j * R - j * (j - 1) / 2 + l - j); };
    T m[kPackBatch];
    auto load = [&](int j0) {
#pragma unroll
        for (int q = 0; q < kPackBatch; ++q) {
            const int j = j0 + q;
            m[q] = clamp ? ld_off(Mt, ok(j) ? cidx(j) : 0u) : (ok(j) ? ld_off(Mt, cidx(j)) : T(0));
        }
    };
    load(0);
    if (l < n) bs[l] = (double)A.x[cl];
    wave_sync();
    double a = 0.0;
#pragma unroll
    for (int q = 0; q < kPackBatch; ++q) a += (!clamp || ok(q) ? (double)m[q] : 0.0) * (q < n ? bs[q] : 0.0);
    if (G > kPackBatch && __ballot(n > kPackBatch)) {
        load(kPackBatch);
#pragma unroll
        for (int q = 0; q < kPackBatch; ++q) {
            const int j = kPackBatch + q;
            a += (!clamp || ok(j) ? (double)m[q] : 0.0) * (j < n ? bs[j] : 0.0);
        }
    }
    if (l < n)
        A.x[cl] = (T)a;
    else if (l < R)
        A.ubuf[uo + l - n] = (T)a;
}

// Backward product of subtree t (t < 0: none) on G lanes, lane l = column l,
// over the rows of the row-major copy (R <= G); v: the G doubles of LDS of
// this lane group.
template <typename T, int G>
__device__ __forceinline__ void sub_bwd_packed(const SubArgs& B, const SolveArgs<T>& A, const T* __restrict__ Mrow,
                                               const int t, const int l, double* v) {
    constexpr bool clamp = sizeof(T) == 4;
    int n = 0, R = 0, cp = 0;
    int64_t mo = 0, ur = 0;
    if (t >= 0) {
        n = B.n[t];
        R = B.R[t];
        cp = B.cptr[t];
        mo = B.moff[t];
        ur = B.urow[t];
    }
    const T* Mt = Mrow + mo;
    // this lane's row value: a subtree column, or one of the root's update rows
    const int xr = l < R ? (l < n ? B.cols[cp + l] : A.rows[ur + l - n]) : 0;
    auto ok = [&](int i) { return i < R && l <= i && l < n; };
    auto ro = [&](int i) { return (uint32_t)(i <= n ? i * (i + 1) / 2 : n * (n + 1) / 2 + (i - n) * n) + l; };
    T mv[kPackBatch];
    auto load = [&](int i0) {
#pragma unroll
        for (int u = 0; u < kPackBatch; ++u) {
            const int i = i0 + u;
            mv[u] = clamp ? ld_off(Mt, ok(i) ? ro(i) : 0u) : (ok(i) ? ld_off(Mt, ro(i)) : T(0));
        }
    };
    load(0);
    if (l < R) v[l] = (double)A.x[xr];
    wave_sync();
    double acc = 0.0;
#pragma unroll
    for (int u = 0; u < kPackBatch; ++u) acc += (!clamp || ok(u) ? (double)mv[u] : 0.0) * (u < R ? v[u] : 0.0);
    if (G > kPackBatch && __ballot(R > kPackBatch)) {
        load(kPackBatch);
#pragma unroll
        for (int u = 0; u < kPackBatch; ++u) {
            const int i = kPackBatch + u;
            acc += (!clamp || ok(i) ? (double)mv[u] : 0.0) * (i < R ? v[i] : 0.0);
        }
    }
    if (l < n) A.x[xr] = (T)acc;  // xr = cols[l] for l < n
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_sub_fwd_pk(SubArgs B, SubPack P, SolveArgs<T> A, const T* __restrict__ M) {
    __shared__ double bsh[kBlock / kWave][kSubCols];
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int wv = blockIdx.x * (kBlock / kWave) + w;
    const int lane = threadIdx.x & (kWave - 1);
    if (wv >= P.nw) return;  // uniform per wave
    const int kind = P.kind[wv];
    if (kind == kWave) {
        sub_fwd_wave<T, 0>(B, A, M, P.sub[4 * wv], lane, bsh[w]);
    } else if (kind == 32) {
        const int g = lane >> 5;
        sub_fwd_packed<T, 32>(B, A, M, P.sub[4 * wv + g], lane & 31, bsh[w] + 32 * g);
    } else {
        const int g = lane >> 4;
        sub_fwd_packed<T, 16>(B, A, M, P.sub[4 * wv + g], lane & 15, bsh[w] + 16 * g);
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_sub_bwd_pk(SubArgs B, SubPack P, SolveArgs<T> A,
                                                       const T* __restrict__ Mrow) {
    __shared__ double vsh[kBlock / kWave][kSubRows];
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int wv = blockIdx.x * (kBlock / kWave) + w;
    const int lane = threadIdx.x & (kWave - 1);
    if (wv >= P.nw) return;  // uniform per wave
    const int kind = P.kind[wv];
    if (kind == kWave) {
        sub_bwd_wave<T, 0>(B, A, Mrow, P.sub[4 * wv], lane, vsh[w]);
    } else if (kind == 32) {
        const int g = lane >> 5;
        sub_bwd_packed<T, 32>(B, A, Mrow, P.sub[4 * wv + g], lane & 31, vsh[w] + 32 * g);
    } else {
        const int g = lane >> 4;
        sub_bwd_packed<T, 16>(B, A, Mrow, P.sub[4 * wv + g], lane & 15, vsh[w] + 16 * g);
    }
}

// Loads through a buffer resource at a wave-uniform byte offset (an SGPR) plus
// a per-lane one, so that a batch of loads shares a single offset register.
// gfx9 descriptor word 3 (32-bit data format); no range limit is used: callers
// keep every access inside the allocation.
__device__ inline __amdgpu_buffer_rsrc_t buf_rsrc(const void* p) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)0xffffffff, 0x00020000);
}
template <typename T>
__device__ inline T buf_ld(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff);
template <>
__device__ inline float buf_ld<float>(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}
template <>
__device__ inline double buf_ld<double>(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0));
}

// The same products with one workgroup per subtree, for the middle level: few
// subtrees (about 1.6 per CU) and most of them near the 64 x 128 limit, so the
// one-wave kernels above spend up to eight dependent batches per subtree with
// little else to hide them.  Here each of the NW waves takes a share of the
// work -- forward the columns w, w + NW, ..., backward the rows w, w + NW, ...
// -- and issues all of its loads at once, ahead of the right-hand-side gather;
// the partial sums then meet in LDS.  The chain per subtree is the gather
// (index, then value) and one reduction.
template <int NW>
constexpr bool sub_split_ok() {
    return NW * kWave >= kSubRows && kSubCols % NW == 0 && kSubRows % NW == 0 && kSubCols == kWave;
}

template <typename T, int NW>
__global__ __launch_bounds__(NW* kWave) void k_sub_fwd_wg(SubArgs B, SolveArgs<T> A, const T* __restrict__ M,
                                                          int t0) {
    static_assert(sub_split_ok<NW>(), "subtree split");
    constexpr int QW = kSubCols / NW;  // columns per wave
    __shared__ double bs[kSubCols];
    __shared__ double part[NW][kSubRows];
    const int t = t0 + blockIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int lane = threadIdx.x & (kWave - 1);
    const int n = B.n[t], R = B.R[t];
    const int32_t* cols = B.cols + B.cptr[t];
    const T* Mt = M + B.moff[t];
    const int cl = threadIdx.x < n ? cols[threadIdx.x] : 0;  // issued first: the gather waits for it alone
    const int64_t uo = B.uoff[t];
    // Column j of the packed column-major M_t starts at j R - j (j - 1) / 2 and
    // holds rows j..R-1, so row i sits at (that start - j) + i: a scalar base per
    // column and the lane as the offset.  Columns past n reread column n - 1;
    // rows outside column j read its neighbours (inside this subtree's copies,
    // or the padding after them) and are discarded in the product.
    const __amdgpu_buffer_rsrc_t rs = buf_rsrc(Mt);
    auto base = [&](int q) {
        const int j = min(w + NW * q, n - 1);
        return (uint32_t)((j * R - j * (j - 1) / 2 - j) * (int)sizeof(T));
    };
    T m0[QW], m1[QW];
#pragma unroll
    for (int q = 0; q < QW; ++q) m0[q] = buf_ld<T>(rs, lane * sizeof(T), base(q));
    if (R > kWave) {
#pragma unroll
        for (int q = 0; q < QW; ++q) m1[q] = buf_ld<T>(rs, (lane + kWave) * sizeof(T), base(q));
    }
    if (threadIdx.x < n) {  // the bottom roots' update rows add into the right-hand side first
        double b = (double)A.x[cl];
        const int c = B.cptr[t] + threadIdx.x;
        for (int q = B.acptr[c]; q < B.acptr[c + 1]; ++q) b += (double)A.ubuf[B.acpos[q]];
        bs[threadIdx.x] = b;
    }
    __syncthreads();
    auto ok = [&](int j, int i) { return j < n && i >= j && i < R; };
    double a0 = 0.0, a1 = 0.0;
#pragma unroll
    for (int q = 0; q < QW; ++q) {
        const int j = w + NW * q;
        const double b = j < n ? bs[j] : 0.0;
        a0 += (ok(j, lane) ? (double)m0[q] : 0.0) * b;
        if (R > kWave) a1 += (ok(j, lane + kWave) ? (double)m1[q] : 0.0) * b;
    }
    part[w][lane] = a0;
    part[w][lane + kWave] = a1;
    __syncthreads();
    const int i = threadIdx.x;
    if (i < R) {
        double y = 0.0;
#pragma unroll
        for (int v = 0; v < NW; ++v) y += part[v][i];
        if (i < n)
            A.x[cl] = (T)y;
        else
            A.ubuf[uo + i - n] = (T)y;
    }
}

template <typename T, int NW>
__global__ __launch_bounds__(NW* kWave) void k_sub_bwd_wg(SubArgs B, SolveArgs<T> A, const T* __restrict__ Mrow,
                                                          int t0) {
    static_assert(sub_split_ok<NW>(), "subtree split");
    constexpr int UW = kSubRows / NW;  // rows per wave
    __shared__ double v[kSubRows];
    __shared__ double part[NW][kSubCols];
    const int t = t0 + blockIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int lane = threadIdx.x & (kWave - 1);
    const int n = B.n[t], R = B.R[t];
    const int32_t* cols = B.cols + B.cptr[t];
    const int32_t* urow = A.rows + B.urow[t];
    const T* Mt = Mrow + B.moff[t];
    const int r = threadIdx.x;
    const int xr = r < R ? (r < n ? cols[r] : urow[r - n]) : 0;  // issued first: the gather waits for it alone
    // row i of the row-major copy starts at sub_row_off(i, n) and holds columns
    // 0..min(i, n - 1): a scalar base per row and the lane as the offset.  Rows
    // past R reread row R - 1; lanes past the row read the next one (or the
    // padding after the copies) and are discarded in the product.
    const __amdgpu_buffer_rsrc_t rs = buf_rsrc(Mt);
    T mv[UW];
#pragma unroll
    for (int u = 0; u < UW; ++u) {
        const int i = min(w + NW * u, R - 1);
        const int ro = i <= n ? i * (i + 1) / 2 : n * (n + 1) / 2 + (i - n) * n;
        mv[u] = buf_ld<T>(rs, lane * sizeof(T), (uint32_t)(ro * (int)sizeof(T)));
    }
    if (r < R) v[r] = (double)A.x[xr];
    __syncthreads();
    double acc = 0.0;
#pragma unroll
    for (int u = 0; u < UW; ++u) {
        const int i = w + NW * u;
        acc += (i < R && lane <= i && lane < n ? (double)mv[u] : 0.0) * (i < R ? v[i] : 0.0);
    }
    part[w][lane] = acc;
    __syncthreads();
    if (r < n) {
        double x = 0.0;
#pragma unroll
        for (int q = 0; q < NW; ++q) x += part[q][r];
        A.x[xr] = (T)x;  // xr = cols[r] for r < n
    }
}

// ---------------------------------------------------------------------------
// Dependency-driven sweeps: one launch covers the whole supernodal tree.
// Block b handles front order[b]; `order` lists every child before its parent
// (forward, factor) or every parent before its children (backward).  A front
// waits until the fronts it depends on have published `epoch` in done[] with
// release semantics (agent scope), works, then publishes its own flag.  The
// blocks a front waits for have lower block indices, so they were dispatched
// earlier and are resident or finished.  The wait is bounded: past spin_limit
// polls the block raises DevStatus::flow_err and proceeds, the host discards
// the sweep and reruns it level by level, so no launch can hang.
// ---------------------------------------------------------------------------
// Polls with relaxed agent-scope loads (served by L2, no cache invalidation per
// poll) and takes one acquire fence once the flag has flipped.
__device__ inline void wait_flag(const uint32_t* flag, uint32_t epoch, int spin_limit, DevStatus* st) {
    int it = 0;
    while (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != epoch) {
        if (++it > spin_limit) {
            st->flow_err = 1;
            break;
        }
        __builtin_amdgcn_s_sleep(2);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

__device__ inline void publish_flag(uint32_t* flag, uint32_t epoch) {
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_store(flag, epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename T>
__global__ __launch_bounds__(kWave) void k_factor_flow(FactArgs<T> A, const int32_t* __restrict__ order,
                                                       uint32_t* done, uint32_t epoch, int spin_limit) {
    extern __shared__ unsigned char smem_raw[];
    T* F = reinterpret_cast<T*>(smem_raw);
    const int s = order[blockIdx.x];
    if (threadIdx.x == 0)
        for (int p = A.child_ptr[s]; p < A.child_ptr[s + 1]; ++p) wait_flag(done + A.child[p], epoch, spin_limit, A.st);
    __syncthreads();
    factor_front<T, kWave>(A, s, F);
    const int c = A.sn_c[s];
    T* Fo = A.F + A.foff[s];
    for (int idx = threadIdx.x; idx < c * c; idx += kWave) Fo[idx] = F[idx];
    publish_flag(done + s, epoch);
}

template <typename T>
__global__ __launch_bounds__(kWave) void k_fwd_flow(SolveArgs<T> A, const int32_t* __restrict__ order,
                                                    uint32_t* done, uint32_t epoch, int spin_limit, DevStatus* st) {
    extern __shared__ unsigned char smem_raw[];
    T* f = reinterpret_cast<T*>(smem_raw);
    const int s = order[blockIdx.x];
    if (threadIdx.x == 0)
        for (int p = A.child_ptr[s]; p < A.child_ptr[s + 1]; ++p) wait_flag(done + A.child[p], epoch, spin_limit, st);
    __syncthreads();
    fwd_front<T>(A, s, f);
    publish_flag(done + s, epoch);
}

template <typename T>
__global__ __launch_bounds__(kWave) void k_bwd_flow(SolveArgs<T> A, const int32_t* __restrict__ order,
                                                    const int32_t* __restrict__ sn_parent, uint32_t* done,
                                                    uint32_t epoch, int spin_limit, DevStatus* st) {
    extern __shared__ unsigned char smem_raw[];
    T* xs = reinterpret_cast<T*>(smem_raw);
    const int s = order[blockIdx.x];
    if (threadIdx.x == 0 && sn_parent[s] >= 0) wait_flag(done + sn_parent[s], epoch, spin_limit, st);
    __syncthreads();
    bwd_front<T>(A, s, xs);
    publish_flag(done + s, epoch);
}

// ---------------------------------------------------------------------------
// Right-hand side b = H^T W r in permuted state order (FP64 accumulation)
// ---------------------------------------------------------------------------
// States are walked in bus-layout order (thread t handles permuted state
// tout[t]) so that the measurement rows it reads are neighbours in memory.
template <typename TO>
__global__ __launch_bounds__(kBlock) void k_rhs(int ns, const int32_t* __restrict__ tptr,
                                                const int32_t* __restrict__ tpos, const int32_t* __restrict__ trow,
                                                const int32_t* __restrict__ tout, const double* __restrict__ hs,
                                                const double* __restrict__ rs, TO* __restrict__ x) {
    int t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= ns) return;
    double acc = 0.0;
    for (int p = tptr[t]; p < tptr[t + 1]; ++p) acc += hs[tpos[p]] * rs[trow[p]];
    x[tout[t]] = (TO)acc;
}

// b = H^T W r from the products k_meas wrote in H^T order: a contiguous
// segmented sum per state (states in bus-layout order, written to x[tout[t]]).
constexpr int kRhsLanes = 8;  // lanes per state in k_rhs_sum

template <typename TO>
__global__ __launch_bounds__(kBlock) void k_rhs_sum(int ns, const int32_t* __restrict__ tptr,
                                                    const int32_t* __restrict__ tout, const double* __restrict__ hr,
                                                    TO* __restrict__ x) {
    // kRhsLanes consecutive lanes read one state's segment together
    const int t = (xcd_block() * kBlock + threadIdx.x) / kRhsLanes;
    const int l = threadIdx.x & (kRhsLanes - 1);
    double acc = 0.0;
    if (t < ns)
        for (int p = tptr[t] + l; p < tptr[t + 1]; p += kRhsLanes) acc += hr[p];
#pragma unroll
    for (int o = kRhsLanes / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, kRhsLanes);
    if (t < ns && l == 0) x[tout[t]] = (TO)acc;
}

__global__ __launch_bounds__(kBlock) void k_gather(int n, const int32_t* __restrict__ idx,
                                                   const double* __restrict__ src, double* __restrict__ dst) {
    int p = blockIdx.x * kBlock + threadIdx.x;
    if (p < n) dst[p] = src[idx[p]];
}

// ---------------------------------------------------------------------------
// Preconditioned conjugate gradients on the FP64 normal equations
// (sqrt(W) H)^T (sqrt(W) H) dx = b, the factor (FP32 in the mixed form) as
// preconditioner.  G p is applied matrix-free as H^T (H p) with the FP64
// Jacobian values, so the Newton step is accurate to FP64 whatever the
// accuracy of the FP32 factor.  All scalars stay on the device: step k reads
// rz[k], pq[k] written by the reductions of the same step.
// ---------------------------------------------------------------------------
constexpr int kMaxCg = 16;
struct CgScalars {
    double rz[kMaxCg + 1];
    double pq[kMaxCg];
};

// t = (sqrt(W) H) p, one measurement row per thread (H columns in permuted order)
__global__ __launch_bounds__(kBlock) void k_hmul(int m, const int32_t* __restrict__ hrow,
                                                 const int32_t* __restrict__ hcolp, const double* __restrict__ hs,
                                                 const double* __restrict__ p, double* __restrict__ t) {
    int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= m) return;
    double acc = 0.0;
    for (int q = hrow[i]; q < hrow[i + 1]; ++q) acc += hs[q] * p[hcolp[q]];
    t[i] = acc;
}

template <typename TA, typename TB>
__global__ __launch_bounds__(kBlock) void k_convert(int ns, const TA* __restrict__ a, TB* __restrict__ b) {
    int k = blockIdx.x * kBlock + threadIdx.x;
    if (k < ns) b[k] = (TB)a[k];
}

__global__ __launch_bounds__(kBlock) void k_dot_part(int ns, const double* __restrict__ a,
                                                     const double* __restrict__ b, double* __restrict__ parts) {
    int k = blockIdx.x * kBlock + threadIdx.x;
    double v = k < ns ? a[k] * b[k] : 0.0;
    __shared__ double sm[kBlock / kWave];
    v = wave_sum(v);
    if ((threadIdx.x & (kWave - 1)) == 0) sm[threadIdx.x / kWave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0.0;
        for (int w = 0; w < kBlock / kWave; ++w) r += sm[w];
        parts[blockIdx.x] = r;
    }
}

__global__ __launch_bounds__(kBlock) void k_dot_finish(int nparts, const double* __restrict__ parts,
                                                       double* __restrict__ out) {
    double v = 0.0;
    for (int p = threadIdx.x; p < nparts; p += kBlock) v += parts[p];
    __shared__ double sm[kBlock / kWave];
    v = wave_sum(v);
    if ((threadIdx.x & (kWave - 1)) == 0) sm[threadIdx.x / kWave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0.0;
        for (int w = 0; w < kBlock / kWave; ++w) r += sm[w];
        *out = r;
    }
}

// x = 0, r = b
__global__ __launch_bounds__(kBlock) void k_cg_init(int ns, const double* __restrict__ b, double* __restrict__ x,
                                                    double* __restrict__ r) {
    int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= ns) return;
    x[k] = 0.0;
    r[k] = b[k];
}

// x += alpha p, r -= alpha q with alpha = rz[k] / pq[k]
__global__ __launch_bounds__(kBlock) void k_cg_xr(int ns, double* __restrict__ x, double* __restrict__ r,
                                                  const double* __restrict__ p, const double* __restrict__ q,
                                                  const CgScalars* __restrict__ cg, int step) {
    int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= ns) return;
    const double pq = cg->pq[step], rz = cg->rz[step];
    const double alpha = (pq > 0.0 && rz > 0.0 && isfinite(rz / pq)) ? rz / pq : 0.0;
    x[k] += alpha * p[k];
    r[k] -= alpha * q[k];
}

// p = z + beta p with beta = rz[k+1] / rz[k]
__global__ __launch_bounds__(kBlock) void k_cg_p(int ns, double* __restrict__ p, const double* __restrict__ z,
                                                 const CgScalars* __restrict__ cg, int step) {
    int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= ns) return;
    const double rz0 = cg->rz[step], rz1 = cg->rz[step + 1];
    const double beta = (rz0 > 0.0 && isfinite(rz1 / rz0)) ? rz1 / rz0 : 0.0;
    p[k] = z[k] + beta * p[k];
}

// ---------------------------------------------------------------------------
// State update (FP64 accumulation), V clamp, max |dx| partials, NaN flag.
// Skipped entirely when the factor raised its pivot flag.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kBlock) void k_update(int n, const int32_t* __restrict__ pth,
                                                   const int32_t* __restrict__ pv, const T* __restrict__ x,
                                                   double* __restrict__ V, double* __restrict__ th, double vmin,
                                                   double vmax, double* __restrict__ mpart, DevStatus* st,
                                                   int skip_on_pivot) {
    const int blk = blockIdx.x;
    const int b = blk * kBlock + threadIdx.x;
    const bool skip = (skip_on_pivot && st->piv != 0) || st->flow_err != 0 || st->stop != 0;
    double mx = 0.0;
    if (b < n && !skip) {
        double dv = (double)x[pv[b]];
        double dt = pth[b] >= 0 ? (double)x[pth[b]] : 0.0;
        if (!isfinite(dv) || !isfinite(dt)) {
            st->nan = 1;
        } else {
            double v = V[b] + dv;
            V[b] = fmin(fmax(v, vmin), vmax);
            th[b] += dt;
            mx = fmax(fabs(dv), fabs(dt));
        }
    }
    __shared__ double sm[kBlock / kWave];
    mx = wave_max(mx);
    if ((threadIdx.x & (kWave - 1)) == 0) sm[threadIdx.x / kWave] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0.0;
        for (int k = 0; k < kBlock / kWave; ++k) r = fmax(r, sm[k]);
        mpart[blk] = r;
    }
}

// Also hands the iteration's status to the host: a copy into the pinned status
// block through its device mapping, so the host waits on the stream and reads
// it without a separate copy (no copy engine, no blit launch per iteration).
// With nj > 0 it also finishes the J / largest-residual partials of a
// measurement pass at the updated state (the pass that closes an iteration), so
// that the status carries J at the state the iteration leaves.  One block of
// kFinish threads; every load (the status flags, the max |dx| and J partials)
// is issued before the first wait.
//
// stop_tol > 0: an iteration whose max |dx| falls below it (or that produced a
// NaN step) sets st->stop, so that an iteration queued behind it changes
// nothing; an iteration that finds st->stop set publishes no status.
constexpr int kFinish = 1024;
__global__ __launch_bounds__(kFinish) void k_maxfinish(int nparts, const double* __restrict__ mpart, DevStatus* st,
                                                       DevStatus* host_out, int nj, const double* __restrict__ jpart,
                                                       const double* __restrict__ rpart, const int* __restrict__ ipart,
                                                       double stop_tol) {
    const int tid = threadIdx.x;
    const int stop = st->stop;
    // thread 0's status fields, loaded with the partials (one round of loads)
    int nan = 0, piv = 0, ferr = 0;
    uint32_t seq = 0;
    double J0 = 0.0, R0 = -1.0;
    int I0 = -1;
    if (tid == 0) {
        nan = st->nan;
        piv = st->piv;
        ferr = st->flow_err;
        seq = st->seq;
        J0 = st->J;
        R0 = st->rmax;
        I0 = st->ridx;
    }
    double mx = 0.0, J = 0.0, R = -1.0;
    int I = -1;
    const int np = max(nparts, nj);
    for (int p0 = tid; p0 < np; p0 += 2 * kFinish) {
        const int p1 = p0 + kFinish;
        const double m0 = p0 < nparts ? mpart[p0] : 0.0, m1 = p1 < nparts ? mpart[p1] : 0.0;
        const double j0 = p0 < nj ? jpart[p0] : 0.0, j1 = p1 < nj ? jpart[p1] : 0.0;
        const double r0 = p0 < nj ? rpart[p0] : -1.0, r1 = p1 < nj ? rpart[p1] : -1.0;
        const int i0 = p0 < nj ? ipart[p0] : -1, i1 = p1 < nj ? ipart[p1] : -1;
        mx = fmax(mx, fmax(m0, m1));
        J += j0 + j1;
        if (r0 > R || (r0 == R && i0 >= 0 && (I < 0 || i0 < I))) {
            R = r0;
            I = i0;
        }
        if (r1 > R || (r1 == R && i1 >= 0 && (I < 0 || i1 < I))) {
            R = r1;
            I = i1;
        }
    }
    if (stop) return;  // uniform: this iteration was queued behind a finished one
    mx = wave_max(mx);
    J = wave_sum(J);
    wave_argmax(R, I);
    __shared__ double smx[kFinish / kWave], sj[kFinish / kWave], sr[kFinish / kWave];
    __shared__ int si[kFinish / kWave];
    const int w = tid / kWave;
    if ((tid & (kWave - 1)) == 0) {
        smx[w] = mx;
        sj[w] = J;
        sr[w] = R;
        si[w] = I;
    }
    __syncthreads();
    if (tid != 0) return;
    double q = 0.0, JJ = 0.0, RR = -1.0;
    int II = -1;
    for (int k = 0; k < kFinish / kWave; ++k) {  // fixed order: deterministic
        q = fmax(q, smx[k]);
        JJ += sj[k];
        if (sr[k] > RR || (sr[k] == RR && si[k] >= 0 && (II < 0 || si[k] < II))) {
            RR = sr[k];
            II = si[k];
        }
    }
    st->maxdx = q;
    if (nj > 0) {
        st->J = JJ;
        st->rmax = RR;
        st->ridx = II;
    }
    const int stopped = stop_tol > 0.0 && !ferr && !piv && (q < stop_tol || nan);
    if (stopped) st->stop = 1;
    if (host_out) {  // two slots, written in turn
        const uint32_t sq = seq + 1;
        st->seq = sq;
        DevStatus* ho = host_out + (sq & 1);
        ho->maxdx = q;
        ho->J = nj > 0 ? JJ : J0;
        ho->rmax = nj > 0 ? RR : R0;
        ho->ridx = nj > 0 ? II : I0;
        ho->nan = nan;
        ho->piv = piv;
        ho->flow_err = ferr;
        ho->stop = stopped;
        // the sequence number last, after the fields are visible to the host
        __threadfence_system();
        __hip_atomic_store(&ho->seq, sq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

__global__ __launch_bounds__(kBlock) void k_flat(int n, const int32_t* __restrict__ bus_type,
                                                 const double* __restrict__ vset, double* __restrict__ V,
                                                 double* __restrict__ th) {
    int b = blockIdx.x * kBlock + threadIdx.x;
    if (b >= n) return;
    V[b] = bus_type[b] != 0 ? vset[b] : 1.0;
    th[b] = 0.0;
}

// The estimate in insertion (natural) bus order, written straight into pinned
// host memory through its device mapping: no copy engine on the read-back.
__global__ __launch_bounds__(kBlock) void k_state_out(int n, const int32_t* __restrict__ lb,
                                                      const double* __restrict__ V, const double* __restrict__ th,
                                                      double* __restrict__ out) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int p = lb[i];
    out[i] = V[p];
    out[n + i] = th[p];
}

__global__ __launch_bounds__(kBlock) void k_scatter_z(int count, const int32_t* __restrict__ idx,
                                                      const double* __restrict__ val, double* __restrict__ z) {
    int q = blockIdx.x * kBlock + threadIdx.x;
    if (q < count) z[idx[q]] = val[q];
}

// z[row of measurement idx[q]] = val[q] (idx in insertion order, mapped on the device)
__global__ __launch_bounds__(kBlock) void k_scatter_z_mapped(int count, const int32_t* __restrict__ idx,
                                                             const double* __restrict__ val,
                                                             const int32_t* __restrict__ lm, double* __restrict__ z) {
    int q = blockIdx.x * kBlock + threadIdx.x;
    if (q < count) z[lm[idx[q]]] = val[q];
}

// natural-order vector <-> permuted solve vector (debug solve only)
template <typename T>
__global__ void k_permute_in(int ns, const int32_t* __restrict__ perm, const double* __restrict__ b,
                             T* __restrict__ x) {
    int k = blockIdx.x * kBlock + threadIdx.x;
    if (k < ns) x[k] = (T)b[perm[k]];
}
template <typename T>
__global__ void k_permute_out(int ns, const int32_t* __restrict__ perm, const T* __restrict__ x,
                              double* __restrict__ dx) {
    int k = blockIdx.x * kBlock + threadIdx.x;
    if (k < ns) dx[perm[k]] = (double)x[k];
}

inline int nblk(int64_t n) { return (int)((n + kBlock - 1) / kBlock); }

// ---------------------------------------------------------------------------
// device buffers
// ---------------------------------------------------------------------------

// Process-wide cache of device blocks.  hipFree waits for the whole device,
// so an engine torn down while another serves requests would stall that
// engine's stream; freed blocks are kept per device and size class (four
// classes per power of two, at most 25% slack) and handed to the next
// allocation of that class instead.  The cache holds the peak of what was in
// use; it is returned to the runtime when the process ends.
class BlockCache {
   public:
    static size_t size_class(size_t bytes) {
        if (bytes <= 4096) return 4096;
        int e = 63 - __builtin_clzll((unsigned long long)bytes);  // 2^e <= bytes < 2^(e+1)
        const size_t q = (size_t)1 << (e - 2);                   // quarter steps
        return (bytes + q - 1) / q * q;
    }
    static void* get(size_t bytes) {
        int dev = 0;
        HIPCHECK(hipGetDevice(&dev));
        const size_t cls = size_class(bytes);
        {
            std::lock_guard<std::mutex> lk(mu());
            auto& fl = free_lists()[{dev, cls}];
            if (!fl.empty()) {
                void* p = fl.back();
                fl.pop_back();
                return p;
            }
        }
        void* p = nullptr;
        HIPCHECK(hipMalloc(&p, cls));
        return p;
    }
    // The releasing instance may still have work queued that reads the block:
    // its stream (set by StreamScope for the duration of an instance call) is
    // drained before the block is offered to any other stream.
    static void put(void* p, size_t bytes) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) return;
        if (release_stream()) (void)hipStreamSynchronize(release_stream());
        std::lock_guard<std::mutex> lk(mu());
        free_lists()[{dev, size_class(bytes)}].push_back(p);
    }

    static hipStream_t& release_stream() {
        static thread_local hipStream_t s = nullptr;
        return s;
    }

   private:
    static std::mutex& mu() {
        static std::mutex m;
        return m;
    }
    static std::map<std::pair<int, size_t>, std::vector<void*>>& free_lists() {
        static auto* f = new std::map<std::pair<int, size_t>, std::vector<void*>>();  // lives to process end
        return *f;
    }
};

// Streams and pinned status blocks kept per device for the life of the
// process.  HIP binds each new stream to one of a few hardware queues in turn
// (GPU_MAX_HW_QUEUES), so engines created and destroyed beside a long-lived
// one would, one creation in a few, land on its queue and serialize their
// kernels with it; a reused stream keeps its queue, and nothing is destroyed or
// freed (which can wait for the whole device) while other engines run.
class EnginePool {
   public:
    static hipStream_t stream() {
        int dev = 0;
        HIPCHECK(hipGetDevice(&dev));
        {
            std::lock_guard<std::mutex> lk(mu());
            auto& fl = streams()[dev];
            if (!fl.empty()) {
                hipStream_t st = fl.back();
                fl.pop_back();
                return st;
            }
        }
        hipStream_t st = nullptr;
        HIPCHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        return st;
    }
    // st must be drained (no queued work)
    static void give_back(hipStream_t st) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) return;
        std::lock_guard<std::mutex> lk(mu());
        streams()[dev].push_back(st);
    }
    static void* pinned(size_t bytes) {
        {
            std::lock_guard<std::mutex> lk(mu());
            auto& fl = pins()[bytes];
            if (!fl.empty()) {
                void* p = fl.back();
                fl.pop_back();
                return p;
            }
        }
        void* p = nullptr;
        HIPCHECK(hipHostMalloc(&p, bytes, hipHostMallocDefault));
        return p;
    }
    static void give_back_pinned(void* p, size_t bytes) {
        std::lock_guard<std::mutex> lk(mu());
        pins()[bytes].push_back(p);
    }

   private:
    static std::mutex& mu() {
        static std::mutex m;
        return m;
    }
    static std::map<int, std::vector<hipStream_t>>& streams() {
        static auto* f = new std::map<int, std::vector<hipStream_t>>();  // lives to process end
        return *f;
    }
    static std::map<size_t, std::vector<void*>>& pins() {
        static auto* f = new std::map<size_t, std::vector<void*>>();
        return *f;
    }
};

// The stream whose queued work may still use the blocks released during an
// instance call.
struct StreamScope {
    hipStream_t prev;
    explicit StreamScope(hipStream_t s) : prev(BlockCache::release_stream()) { BlockCache::release_stream() = s; }
    ~StreamScope() { BlockCache::release_stream() = prev; }
};

template <typename T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    void alloc(size_t cnt, Ledger& bytes) {
        if (bytes.frozen) throw NoAllocError("the request needs a device allocation (allocation-free solve)");
        release(bytes);
        n = cnt;
        if (cnt) {
            p = static_cast<T*>(BlockCache::get(cnt * sizeof(T)));
            bytes.bytes += cnt * sizeof(T);
            ++bytes.epoch;
        }
    }
    void release(Ledger& bytes) {
        if (p && bytes.frozen) throw NoAllocError("the request would release a device buffer (allocation-free solve)");
        if (p) {
            BlockCache::put(p, n * sizeof(T));
            bytes.bytes -= n * sizeof(T);
            ++bytes.epoch;
        }
        p = nullptr;
        n = 0;
    }
    void upload(const T* h, size_t cnt, hipStream_t s) {
        if (cnt) HIPCHECK(hipMemcpyAsync(p, h, cnt * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void upload(const std::vector<T>& v, Ledger& bytes, hipStream_t s) {
        if (n != v.size()) alloc(v.size(), bytes);
        upload(v.data(), v.size(), s);
    }
};

// Sizes of the partitioned-inverse buffers (zero when that solve is off).
struct PiSizes {
    int64_t nblock = 0, nq = 0, msub = 0;
    int ntc = 0;
};

template <typename T>
struct PrecBufs {
    DBuf<T> Gval, F, x, ubuf, M;
    DBuf<T> pinv, pinv_rows, ctop, ytop;  // L_TT^-1 blocks, row-grouped copy, top vectors
    DBuf<T> msub;                         // dense solve matrices of the bottom subtrees
    bool ready = false;
    void alloc(const Symbolic& S, Ledger& bytes, const PiSizes& pz) {
        Gval.alloc(std::max<int64_t>(1, S.nnz_g), bytes);
        F.alloc(std::max<int64_t>(1, S.ftotal), bytes);
        x.alloc(std::max(1, S.ns), bytes);
        ubuf.alloc(std::max<int64_t>(1, S.utotal), bytes);
        M.alloc(std::max<int64_t>(1, S.mtotal), bytes);
        alloc_pi(pz, bytes);
        ready = true;
    }
    void alloc_pi(const PiSizes& pz, Ledger& bytes) {
        pinv.alloc(std::max<int64_t>(1, pz.nblock), bytes);
        pinv_rows.alloc(std::max<int64_t>(1, pz.nq), bytes);
        ctop.alloc(std::max(1, pz.ntc), bytes);
        ytop.alloc(std::max(1, pz.ntc), bytes);
        // column-major, then row-major; the workgroup products read up to a
        // row's width (kSubCols) past the last row of the last subtree
        msub.alloc(2 * pz.msub + kSubCols, bytes);
    }
    void release(Ledger& bytes) {
        for (DBuf<T>* b : {&Gval, &F, &x, &ubuf, &M, &pinv, &pinv_rows, &ctop, &ytop, &msub}) b->release(bytes);
        ready = false;
    }
};

}  // namespace

// ---------------------------------------------------------------------------
// Sweep plan for the task sweeps (host).  Fronts of subtrees with at most
// `cut` fronts form bottom tasks; the remaining top of the tree is walked level
// by level (one launch per level) while a level holds more than `wcut` fronts,
// and the rest (the narrow top) is one single-workgroup task.  A small cost
// model (per-front latency, per-launch overhead) picks cut and wcut.
// ---------------------------------------------------------------------------
struct SweepPlan {
    std::vector<int32_t> fronts, lvl_ptr, lvl_off, mid;
    std::vector<int> mid_off, mid_cnt;
    int n_bottom = 0;
    bool has_tail = false;
    int cut = 0, wcut = 0;
    double cost_us = 0;
    // level lists kept for the LDS-resident solve plan
    std::vector<std::vector<std::vector<int32_t>>> bottom_levels;
    std::vector<std::vector<int32_t>> tail_levels;
};

constexpr double kFrontUs = 2.0, kLaunchUs = 8.0;

static void append_task(SweepPlan& P, const std::vector<std::vector<int32_t>>& levels) {
    P.lvl_off.push_back((int32_t)P.lvl_ptr.size());
    for (const auto& L : levels) {
        P.lvl_ptr.push_back((int32_t)P.fronts.size());
        P.fronts.insert(P.fronts.end(), L.begin(), L.end());
    }
    P.lvl_ptr.push_back((int32_t)P.fronts.size());
}

static SweepPlan make_plan(const Symbolic& S, int cut, int wcut, int bottom_waves, int tail_waves, bool build) {
    const int nsn = S.nsn;
    SweepPlan P;
    P.cut = cut;
    P.wcut = wcut;
    std::vector<int32_t> size(nsn, 1);
    for (int s = 0; s < nsn; ++s)
        if (S.sn_parent[s] >= 0) size[S.sn_parent[s]] += size[s];
    auto inside = [&](int s) { return size[s] <= cut; };
    // bottom tasks: maximal subtrees with <= cut fronts, fronts by level
    std::vector<int32_t> root_of(nsn, -1);
    std::vector<int32_t> roots;
    for (int s = nsn - 1; s >= 0; --s) {  // parents have larger indices
        if (!inside(s)) continue;
        int p = S.sn_parent[s];
        if (p < 0 || !inside(p)) {
            root_of[s] = (int32_t)roots.size();
            roots.push_back(s);
        } else {
            root_of[s] = root_of[p];
        }
    }
    std::vector<std::vector<std::vector<int32_t>>> tl(roots.size());
    for (size_t t = 0; t < roots.size(); ++t) tl[t].resize(S.sn_level[roots[t]] + 1);
    for (int s = 0; s < nsn; ++s)
        if (root_of[s] >= 0) tl[root_of[s]][S.sn_level[s]].push_back(s);
    // the fronts of a level go to the waves in turn: largest fronts first, so
    // that the waves' shares of a level end close together
    auto by_size = [&](std::vector<int32_t>& L) {
        std::stable_sort(L.begin(), L.end(), [&](int a, int b) { return S.sn_c[a] > S.sn_c[b]; });
    };
    for (auto& T : tl)
        for (auto& L : T) by_size(L);
    double costA = 0;
    std::vector<std::pair<double, int>> tcost(roots.size());
    for (size_t t = 0; t < roots.size(); ++t) {
        double c = 0;
        for (const auto& L : tl[t]) c += std::ceil((double)L.size() / bottom_waves) * kFrontUs;
        costA = std::max(costA, c);
        tcost[t] = {-c, (int)t};
    }
    // top tree levels
    std::vector<int32_t> htop(nsn, 0);
    int H = 0;
    for (int s = 0; s < nsn; ++s) {
        if (inside(s)) continue;
        H = std::max(H, htop[s] + 1);
        int p = S.sn_parent[s];
        if (p >= 0) htop[p] = std::max(htop[p], htop[s] + 1);
    }
    std::vector<std::vector<int32_t>> top(H);
    for (int s = 0; s < nsn; ++s)
        if (!inside(s)) top[htop[s]].push_back(s);
    for (auto& L : top) by_size(L);
    int Lstar = H;
    while (Lstar > 0 && (int)top[Lstar - 1].size() <= wcut) --Lstar;
    double costB = Lstar * (kLaunchUs + kFrontUs);
    double costC = 0;
    for (int l = Lstar; l < H; ++l) costC += std::ceil((double)top[l].size() / tail_waves) * kFrontUs;
    P.cost_us = (roots.empty() ? 0 : costA + kLaunchUs) + costB + (Lstar < H ? costC + kLaunchUs : 0);
    if (!build) return P;
    // bottom tasks longest first: the workgroups dispatched last are the short ones
    std::sort(tcost.begin(), tcost.end());
    std::vector<std::vector<std::vector<int32_t>>> tls;
    tls.reserve(roots.size());
    for (const auto& tc : tcost) tls.push_back(std::move(tl[tc.second]));
    for (const auto& L : tls) append_task(P, L);
    P.n_bottom = (int)roots.size();
    P.bottom_levels = std::move(tls);
    for (int l = 0; l < Lstar; ++l) {
        P.mid_off.push_back((int)P.mid.size());
        P.mid_cnt.push_back((int)top[l].size());
        P.mid.insert(P.mid.end(), top[l].begin(), top[l].end());
    }
    if (Lstar < H) {
        std::vector<std::vector<int32_t>> tail(top.begin() + Lstar, top.end());
        append_task(P, tail);
        P.has_tail = true;
        P.tail_levels = tail;
    }
    P.lvl_off.push_back((int32_t)P.lvl_ptr.size());
    return P;
}


// LDS-resident solve plan built on a SweepPlan: every bottom subtree is one
// task; the narrow top is cut into chunks of consecutive levels that fit the
// LDS budget.  Returns false when a bottom task does not fit.
struct LdsPlan {
    std::vector<int32_t> th, meta, lvl, ch, rm, src;
    int n_bottom = 0;
    std::vector<int> tail_tasks;
    std::vector<size_t> tail_bytes;
    size_t bottom_bytes = 0;
    std::vector<int> mid_off, mid_cnt;
    std::vector<int32_t> mid;
    double cost_us = 0;
    int cut = 0, wcut = 0;
};

constexpr double kLdsLevelUs = 0.6, kLdsFixedUs = 6.0, kMidLevelUs = 10.0;

static size_t lds_task_bytes(const Symbolic& S, const std::vector<std::vector<int32_t>>& levels, size_t elem,
                             int64_t* nfronts = nullptr) {
    int64_t nf = 0, sumL = 0, sumF = 0, nch = 0, nrm = 0, nsrc = 0;
    for (const auto& L : levels)
        for (int s : L) {
            ++nf;
            sumL += (int64_t)S.sn_c[s] * S.sn_k[s];
            sumF += S.sn_c[s];
            nch += S.sn_child_ptr[s + 1] - S.sn_child_ptr[s];
            nrm += S.sn_c[s] - S.sn_k[s];
            nsrc += S.sn_c[s] - S.sn_k[s];
        }
    if (nfronts) *nfronts = nf;
    return (size_t)(sumL + sumF) * elem + 4 * (size_t)(nf * kFM + levels.size() + 1 + nch + nrm + nsrc + nf);
}

static void add_lds_task(LdsPlan& P, const Symbolic& S, const std::vector<int32_t>& snode_of_col,
                         std::vector<int32_t>& local, const std::vector<std::vector<int32_t>>& levels) {
    const int32_t meta_off = (int32_t)P.meta.size(), lvl_off = (int32_t)P.lvl.size(), ch_off = (int32_t)P.ch.size(),
                  rm_off = (int32_t)P.rm.size(), src_off = (int32_t)P.src.size();
    std::vector<int32_t> order;
    for (const auto& L : levels) {
        P.lvl.push_back((int32_t)order.size());
        for (int s : L) {
            local[s] = (int32_t)order.size();
            order.push_back(s);
        }
    }
    P.lvl.push_back((int32_t)order.size());
    std::vector<int32_t> Foff(order.size());
    int32_t sumL = 0, sumF = 0;
    for (size_t lf = 0; lf < order.size(); ++lf) {
        Foff[lf] = sumF;
        sumF += S.sn_c[order[lf]];
    }
    for (size_t lf = 0; lf < order.size(); ++lf) {
        const int s = order[lf], c = S.sn_c[s], k = S.sn_k[s], first = S.sn_first[s];
        int32_t m[kFM];
        m[FM_S] = s;
        m[FM_C] = c;
        m[FM_K] = k;
        m[FM_FIRST] = first;
        m[FM_L] = sumL;
        sumL += c * k;
        m[FM_F] = Foff[lf];
        m[FM_CH] = (int32_t)P.ch.size() - ch_off;
        m[FM_NCH] = S.sn_child_ptr[s + 1] - S.sn_child_ptr[s];
        for (int q = S.sn_child_ptr[s]; q < S.sn_child_ptr[s + 1]; ++q) {
            const int chd = S.sn_child[q];
            P.ch.push_back(local[chd] >= 0 ? local[chd] : -chd - 1);
        }
        const int par = S.sn_parent[s];
        m[FM_PAR] = (par >= 0 && local[par] >= 0) ? local[par] : -1;
        // the flag-synchronised sweeps need every parent listed after its children
        if (m[FM_PAR] >= 0 && m[FM_PAR] <= (int32_t)lf) throw std::logic_error("task fronts out of level order");
        if (par >= 0 && local[par] >= 0) {
            m[FM_RM] = (int32_t)P.rm.size() - rm_off;
            m[FM_EXTPAR] = 0;
            P.rm.insert(P.rm.end(), S.rmap.begin() + S.sn_rmoff[s], S.rmap.begin() + S.sn_rmoff[s] + (c - k));
        } else {
            m[FM_RM] = -1;
            m[FM_EXTPAR] = c - k > 0 ? 1 : 0;
        }
        m[FM_SRC] = (int32_t)P.src.size() - src_off;
        for (int i = k; i < c; ++i) {
            const int r = S.sn_rows[S.sn_rows_off[s] + i];
            const int owner = snode_of_col[r];
            P.src.push_back(local[owner] >= 0 ? Foff[local[owner]] + (r - S.sn_first[owner]) : -r - 1);
        }
        P.meta.insert(P.meta.end(), m, m + kFM);
    }
    const int32_t hdr[kTH] = {(int32_t)order.size(), (int32_t)levels.size(), meta_off, lvl_off, ch_off,
                              (int32_t)P.ch.size() - ch_off, rm_off, (int32_t)P.rm.size() - rm_off, src_off,
                              (int32_t)P.src.size() - src_off, sumL, sumF};
    P.th.insert(P.th.end(), hdr, hdr + kTH);
    for (int s : order) local[s] = -1;
}

static bool make_lds_plan(const Symbolic& S, const SweepPlan& SP, size_t elem, size_t budget, int nwb, int nwt,
                          bool build, LdsPlan& P) {
    P = LdsPlan();
    P.cut = SP.cut;
    P.wcut = SP.wcut;
    double costA = 0;
    for (const auto& T : SP.bottom_levels) {
        size_t b = lds_task_bytes(S, T, elem);
        if (b > budget) return false;
        P.bottom_bytes = std::max(P.bottom_bytes, b);
        double c = kLdsFixedUs;
        for (const auto& L : T) c += std::ceil((double)L.size() / nwb) * kLdsLevelUs;
        costA = std::max(costA, c);
    }
    // tail chunks of consecutive levels within the budget
    std::vector<std::vector<std::vector<int32_t>>> chunks;
    std::vector<std::vector<int32_t>> cur;
    double costC = 0;
    for (const auto& L : SP.tail_levels) {
        cur.push_back(L);
        if (lds_task_bytes(S, cur, elem) > budget) {
            cur.pop_back();
            if (cur.empty()) return false;
            chunks.push_back(cur);
            cur.assign(1, L);
            if (lds_task_bytes(S, cur, elem) > budget) return false;
        }
    }
    if (!cur.empty()) chunks.push_back(cur);
    for (const auto& C : chunks) {
        costC += kLaunchUs + kLdsFixedUs;
        for (const auto& L : C) costC += std::ceil((double)L.size() / nwt) * kLdsLevelUs;
    }
    P.cost_us = (SP.bottom_levels.empty() ? 0 : kLaunchUs + costA) + SP.mid_off.size() * kMidLevelUs + costC;
    if (!build) return true;
    std::vector<int32_t> snode_of_col(S.ns), local(S.nsn, -1);
    for (int sn = 0; sn < S.nsn; ++sn)
        for (int j = 0; j < S.sn_k[sn]; ++j) snode_of_col[S.sn_first[sn] + j] = sn;
    for (const auto& T : SP.bottom_levels) add_lds_task(P, S, snode_of_col, local, T);
    P.n_bottom = (int)SP.bottom_levels.size();
    for (const auto& C : chunks) {
        P.tail_tasks.push_back((int)P.th.size() / kTH);
        P.tail_bytes.push_back(lds_task_bytes(S, C, elem));
        add_lds_task(P, S, snode_of_col, local, C);
    }
    P.mid = SP.mid;
    P.mid_off = SP.mid_off;
    P.mid_cnt = SP.mid_cnt;
    return true;
}

// Partitioned-inverse plan (host): bottom subtrees of at most `cut` fronts, the
// rest is the top T.  Block, path, forward-product and assembly index lists.
struct PiPlan {
    int cut = 0;
    int ntop = 0, ntc = 0, nseg = 0, slice = 0;
    int64_t nblock = 0, nq = 0;
    std::vector<int32_t> sn, plen, ploc, tcol, colsn, src, segptr, cptr, parent, segdst, mrows, cplen, cjj, ufirst,
        uk, bunit, upath, upoff;
    std::vector<int64_t> boff, poff, pos, segq, segc, cpos, cboff, cpoff;
    // dense bottom subtrees (see SubArgs)
    bool dense = false;
    int nsub = 0, nsub0 = 0;  // subtrees; of them bottom ones (listed first)
    int64_t msub = 0;
    std::vector<int32_t> sroot, sn_n, sR, fptr, fsn, fdcol, drow, scptr, scols, bins, gsrc, gdst, grs;
    std::vector<int64_t> geptr, gbase, gmo;
    std::vector<uint64_t> gmask;
    // middle columns; and per subtree column (in scols order) the bottom roots'
    // update positions that add into it (empty for bottom subtree columns)
    std::vector<int32_t> mcol, acptr;
    std::vector<int64_t> acpos;
    std::vector<int64_t> fdoff, smoff;
    // per subtree: its root's update rows in the update buffer and in the row lists
    std::vector<int64_t> suoff, surow;
    int bin_off[kSubBins + 1] = {0, 0, 0, 0, 0};
    size_t bin_bytes[kSubBins] = {0, 0, 0, 0};
    // the bottom-subtree products' wave schedule (SubPack)
    std::vector<int32_t> pkind, psub;
};

// The wave schedule of the bottom-subtree products: subtrees of at most 16
// rows four to a wave, of at most 32 rows two to a wave, the others one to a
// wave; the waves largest first (by the entries of their largest subtree).
static void make_sub_pack(PiPlan& P) {
    struct Wave {
        int64_t cost;
        int kind;
        int32_t sub[4];
    };
    std::vector<Wave> waves;
    Wave open[2] = {{0, 16, {-1, -1, -1, -1}}, {0, 32, {-1, -1, -1, -1}}};
    int fill[2] = {0, 0};
    const int cap[2] = {4, 2};
    for (int t = 0; t < P.nsub0; ++t) {  // the plan lists them largest first
        const int R = P.sR[t];
        const int64_t e = sub_col_off(P.sn_n[t], R);
        const int c = R <= 16 ? 0 : (R <= 32 ? 1 : -1);
        if (c < 0) {
            waves.push_back({e, kWave, {t, -1, -1, -1}});
            continue;
        }
        if (fill[c] == 0) open[c].cost = e;
        open[c].sub[fill[c]++] = t;
        if (fill[c] == cap[c]) {
            waves.push_back(open[c]);
            open[c] = {0, c ? 32 : 16, {-1, -1, -1, -1}};
            fill[c] = 0;
        }
    }
    for (int c = 0; c < 2; ++c)
        if (fill[c]) waves.push_back(open[c]);
    std::stable_sort(waves.begin(), waves.end(), [](const Wave& a, const Wave& b) { return a.cost > b.cost; });
    P.pkind.clear();
    P.psub.clear();
    for (const Wave& w : waves) {
        P.pkind.push_back(w.kind);
        P.psub.insert(P.psub.end(), w.sub, w.sub + 4);
    }
}

// Bottom subtrees as dense solve matrices: for every subtree root (a front
// outside the top whose parent is in the top, or a root of the tree) list its
// fronts children first, its columns in that order, and the dense row of every
// front row (a subtree column, or a root update row placed after the columns).
static void make_sub_plan(const Symbolic& S, const std::vector<int8_t>& lev, PiPlan& P) {
    const int nsn = S.nsn;
    std::vector<int32_t> dpos(S.ns, -1), stack, fronts;
    P.fptr.push_back(0);
    P.scptr.push_back(0);
    // the bottom subtrees (level 0) first, then the middle ones (level 1): a
    // subtree is its root's descendants of the same level.  Within a level the
    // largest subtrees come first (most matrix entries), so that the waves of
    // the solve products that start last are the short ones.
    for (int level = 0; level < 2; ++level) {
    std::vector<std::pair<int64_t, int>> order;
    for (int r = 0; r < nsn; ++r) {
        const int par = S.sn_parent[r];
        if (lev[r] != level || (par >= 0 && lev[par] == level)) continue;
        int64_t n = 0;
        stack.assign(1, r);
        while (!stack.empty()) {
            const int s = stack.back();
            stack.pop_back();
            n += S.sn_k[s];
            for (int q = S.sn_child_ptr[s]; q < S.sn_child_ptr[s + 1]; ++q)
                if (lev[S.sn_child[q]] == level) stack.push_back(S.sn_child[q]);
        }
        const int64_t R = n + S.sn_c[r] - S.sn_k[r];
        order.emplace_back(-(n * R - n * (n - 1) / 2), r);
    }
    std::sort(order.begin(), order.end());
    for (const auto& o : order) {
        const int r = o.second;
        if (level == 0) P.nsub0 = (int)P.sroot.size() + 1;
        fronts.clear();
        stack.assign(1, r);
        while (!stack.empty()) {
            const int s = stack.back();
            stack.pop_back();
            fronts.push_back(s);
            for (int q = S.sn_child_ptr[s]; q < S.sn_child_ptr[s + 1]; ++q)
                if (lev[S.sn_child[q]] == level) stack.push_back(S.sn_child[q]);
        }
        std::sort(fronts.begin(), fronts.end());  // children carry smaller ids than their parents
        int n = 0;
        for (int s : fronts) {
            P.fsn.push_back(s);
            P.fdcol.push_back(n);
            for (int j = 0; j < S.sn_k[s]; ++j) {
                dpos[S.sn_first[s] + j] = n + j;
                P.scols.push_back(S.sn_first[s] + j);
            }
            n += S.sn_k[s];
        }
        const int kr = S.sn_k[r], u = S.sn_c[r] - kr, R = n + u;
        if (n > kSubCols || R > kSubRows) throw std::logic_error("dense subtree larger than its limits");
        for (int a = 0; a < u; ++a) dpos[S.sn_rows[S.sn_rows_off[r] + kr + a]] = n + a;
        for (int s : fronts) {
            P.fdoff.push_back((int64_t)P.drow.size());
            for (int i = 0; i < S.sn_c[s]; ++i) {
                const int d = dpos[S.sn_rows[S.sn_rows_off[s] + i]];
                if (d < 0) throw std::logic_error("dense subtree: front row outside the subtree and its root rows");
                P.drow.push_back(d);
            }
        }
        // the gather's entry list (column j of front s, rows j .. c - 1) and
        // the rows of L as CSR: per row its column mask and first entry
        {
            const int64_t base = S.sn_foff[fronts[0]];  // the smallest front id: the lowest offset
            if (P.geptr.empty()) P.geptr.push_back(0);
            P.gbase.push_back(base);
            P.gmo.push_back((int64_t)P.gmask.size());
            const size_t f0 = P.fsn.size() - fronts.size();
            std::vector<uint64_t> mask(R, 0);
            for (int pass = 0; pass < 2; ++pass) {
                std::vector<int32_t> rs(R + 1, 0);
                for (int r = 0; r < R; ++r) rs[r + 1] = rs[r] + __builtin_popcountll(mask[r]);
                for (size_t q = 0; q < fronts.size(); ++q) {
                    const int s = fronts[q], c = S.sn_c[s], k = S.sn_k[s], dc = P.fdcol[f0 + q];
                    const int64_t fo = S.sn_foff[s] - base;
                    const int64_t dr = P.fdoff[f0 + q];
                    if (fo + (int64_t)c * c > INT32_MAX) throw std::logic_error("dense subtree: fronts too far apart");
                    for (int j = 0; j < k; ++j)
                        for (int i = j; i < c; ++i) {
                            const int r = P.drow[dr + i], col = dc + j;
                            if (pass == 0) {
                                mask[r] |= 1ull << col;
                            } else {
                                P.gsrc.push_back((int32_t)(fo + (int64_t)j * c + i));
                                P.gdst.push_back(rs[r] + __builtin_popcountll(mask[r] & ((1ull << col) - 1)));
                            }
                        }
                }
                if (pass == 1) {
                    P.gmask.insert(P.gmask.end(), mask.begin(), mask.end());
                    P.grs.insert(P.grs.end(), rs.begin(), rs.end());
                    for (int r = 0; r < n; ++r)  // the diagonal closes each row of L_BB
                        if (!(mask[r] >> r & 1) || (mask[r] >> r) != 1)
                            throw std::logic_error("dense subtree: a row of L without its diagonal last");
                }
            }
            P.geptr.push_back((int64_t)P.gsrc.size());
        }
        for (int s : fronts)
            for (int j = 0; j < S.sn_k[s]; ++j) dpos[S.sn_first[s] + j] = -1;
        for (int a = 0; a < u; ++a) dpos[S.sn_rows[S.sn_rows_off[r] + kr + a]] = -1;
        P.sroot.push_back(r);
        P.suoff.push_back(S.sn_uoff[r]);
        P.surow.push_back((int64_t)S.sn_rows_off[r] + kr);
        P.sn_n.push_back(n);
        P.sR.push_back(R);
        P.smoff.push_back(P.msub);
        P.msub += sub_col_off(n, R);
        P.fptr.push_back((int32_t)P.fsn.size());
        P.scptr.push_back((int32_t)P.scols.size());
    }
    }
    P.nsub = (int)P.sroot.size();
    // launches of the subtree-matrix kernel k_sub_m_rows grouped by the LDS
    // they need (L's entries, packed X, row masks and starts: at most 82 KiB for
    // a dense 128 x 64, within the 160 KiB a gfx950 workgroup may hold)
    const size_t lim[3] = {16384, 32768, 65536};
    std::vector<int32_t> bin[kSubBins];
    for (int t = 0; t < P.nsub; ++t) {
        const size_t n = P.sn_n[t], R = P.sR[t];
        const size_t b = sub_m_lds(P.geptr[t + 1] - P.geptr[t], (int)n, (int)R);
        const int q = b <= lim[0] ? 0 : (b <= lim[1] ? 1 : (b <= lim[2] ? 2 : 3));
        bin[q].push_back(t);
        P.bin_bytes[q] = std::max(P.bin_bytes[q], b);
    }
    for (int q = 0; q < kSubBins; ++q) {  // largest first within a bin
        std::stable_sort(bin[q].begin(), bin[q].end(), [&](int a, int b) {
            return (int64_t)P.sR[a] * P.sn_n[a] > (int64_t)P.sR[b] * P.sn_n[b];
        });
        P.bin_off[q] = (int)P.bins.size();
        P.bins.insert(P.bins.end(), bin[q].begin(), bin[q].end());
    }
    P.bin_off[kSubBins] = (int)P.bins.size();
    P.dense = true;
    make_sub_pack(P);
}

constexpr int kPiSeg = kPiSegMax;  // entries per forward-product segment
constexpr int kPiCutDefault = 32;  // fronts per bottom subtree

static PiPlan make_pi_plan(const Symbolic& S, int cut, bool dense, bool mid) {
    PiPlan P;
    P.cut = cut;
    const int nsn = S.nsn;
    std::vector<int32_t> size(nsn, 1), ncol(S.sn_k.begin(), S.sn_k.end());
    for (int s = 0; s < nsn; ++s)
        if (S.sn_parent[s] >= 0) {
            size[S.sn_parent[s]] += size[s];
            ncol[S.sn_parent[s]] += ncol[s];
        }
    // a front is in the top when its subtree is too large to be a bottom task;
    // every bound grows towards the root, so the top is closed under parents
    auto in_top = [&](int s) {
        if (size[s] > cut) return true;
        return dense && (ncol[s] > kSubCols || ncol[s] + S.sn_c[s] - S.sn_k[s] > kSubRows);
    };
    // level of each front: 0 bottom subtree, 1 middle subtree (dense subtrees
    // of the forest left above the bottom ones, same limits), 2 top
    std::vector<int8_t> lev(nsn, 0);
    for (int s = 0; s < nsn; ++s) lev[s] = in_top(s) ? 2 : 0;
    if (dense && mid) {
        std::vector<int32_t> nc2(nsn, 0);
        for (int s = 0; s < nsn; ++s) {
            if (lev[s] != 2) continue;
            nc2[s] += S.sn_k[s];
            const int p = S.sn_parent[s];
            if (p >= 0 && lev[p] == 2) nc2[p] += nc2[s];
        }
        // both bounds grow towards the root: the middle subtrees are closed
        // downwards inside the top forest, what stays at level 2 upwards
        for (int s = 0; s < nsn; ++s)
            if (lev[s] == 2 && nc2[s] <= kSubCols && nc2[s] + S.sn_c[s] - S.sn_k[s] <= kSubRows) lev[s] = 1;
    }
    std::vector<int32_t> tloc(S.ns, -1), tsl(nsn, -1), mloc(S.ns, -1);
    for (int s = 0; s < nsn; ++s)
        if (lev[s] == 1)
            for (int j = 0; j < S.sn_k[s]; ++j) {
                mloc[S.sn_first[s] + j] = (int32_t)P.mcol.size();
                P.mcol.push_back(S.sn_first[s] + j);
            }
    // the top fronts in a postorder of the top forest (children in ascending
    // order): the descendants of a front, and so their columns, are then one
    // contiguous range just before its own -- every row of L_TT^-1 reads one
    // contiguous range of c_T in the forward product
    std::vector<int32_t> top_order;
    {
        std::vector<int32_t> cptr_t(nsn + 1, 0), ch_t, roots, fill_t;
        for (int s = 0; s < nsn; ++s)
            if (lev[s] == 2 && S.sn_parent[s] >= 0) ++cptr_t[S.sn_parent[s] + 1];
        for (int s = 0; s < nsn; ++s) cptr_t[s + 1] += cptr_t[s];
        ch_t.resize(cptr_t[nsn]);
        fill_t.assign(cptr_t.begin(), cptr_t.end() - 1);
        for (int s = 0; s < nsn; ++s) {
            if (lev[s] != 2) continue;
            if (S.sn_parent[s] >= 0)
                ch_t[fill_t[S.sn_parent[s]]++] = s;  // ascending within each parent
            else
                roots.push_back(s);
        }
        std::vector<std::pair<int32_t, int32_t>> st;  // (front, next child)
        for (int r : roots) {
            st.emplace_back(r, cptr_t[r]);
            while (!st.empty()) {
                auto& f = st.back();
                if (f.second < cptr_t[f.first + 1]) {
                    const int c = ch_t[f.second++];
                    st.emplace_back(c, cptr_t[c]);
                } else {
                    top_order.push_back(f.first);
                    st.pop_back();
                }
            }
        }
    }
    int cmax = 1, kmax = 1;
    for (int s : top_order) {
        tsl[s] = (int32_t)P.sn.size();
        P.sn.push_back(s);
        for (int j = 0; j < S.sn_k[s]; ++j) {
            tloc[S.sn_first[s] + j] = (int32_t)P.tcol.size();
            P.tcol.push_back(S.sn_first[s] + j);
            P.colsn.push_back(tsl[s]);
        }
        cmax = std::max(cmax, S.sn_c[s]);
        kmax = std::max(kmax, S.sn_k[s]);
    }
    P.ntop = (int)P.sn.size();
    P.ntc = (int)P.tcol.size();
    P.slice = cmax * kmax;
    P.parent = S.sn_parent;
    // blocks and path rows
    P.plen.resize(P.ntop);
    P.boff.resize(P.ntop);
    P.poff.resize(P.ntop);
    P.ufirst.resize(P.ntop);
    P.uk.resize(P.ntop);
    P.upoff.assign(1, 0);
    for (int u = 0; u < P.ntop; ++u) {
        const int s0 = P.sn[u];
        P.ufirst[u] = S.sn_first[s0];
        P.uk[u] = S.sn_k[s0];
        for (int p = s0; p >= 0; p = S.sn_parent[p]) P.upath.push_back(p);
        if (P.upath.size() > (size_t)INT32_MAX) throw std::logic_error("partitioned inverse: top paths too long");
        P.upoff.push_back((int32_t)P.upath.size());
        P.boff[u] = P.nblock;
        P.poff[u] = (int64_t)P.ploc.size();
        int len = 0;
        for (int p = s0; p >= 0; p = S.sn_parent[p])
            for (int j = 0; j < S.sn_k[p]; ++j, ++len) P.ploc.push_back(tloc[S.sn_first[p] + j]);
        P.plen[u] = len;
        for (int j = 0; j < S.sn_k[s0]; ++j) {
            P.cboff.push_back(P.boff[u] + (int64_t)len * j);
            P.cpoff.push_back(P.poff[u]);
            P.cplen.push_back(len);
            P.cjj.push_back(j);
        }
        P.nblock += (int64_t)len * S.sn_k[s0];
    }
    // backward-product units: kPiBwdCols columns of one supernode, longest path first
    for (int u = 0; u < P.ntop; ++u) {
        if (u >= (1 << (31 - kPiUnitShift)) || P.uk[u] > (1 << kPiUnitShift))
            throw std::logic_error("partitioned inverse: top too large for the backward-product units");
        for (int j0 = 0; j0 < P.uk[u]; j0 += kPiBwdCols) P.bunit.push_back(u << kPiUnitShift | j0);
    }
    std::stable_sort(P.bunit.begin(), P.bunit.end(), [&](int32_t a, int32_t b) {
        return P.plen[a >> kPiUnitShift] > P.plen[b >> kPiUnitShift];
    });
    // forward product grouped by row: entry (u, r, jj) lands in row ploc[r]
    std::vector<int64_t> rcnt(P.ntc + 1, 0);
    for (int u = 0; u < P.ntop; ++u) {
        const int ks = S.sn_k[P.sn[u]];
        for (int r = 0; r < P.plen[u]; ++r) rcnt[P.ploc[P.poff[u] + r] + 1] += std::min(r + 1, ks);
    }
    for (int t = 0; t < P.ntc; ++t) rcnt[t + 1] += rcnt[t];
    P.nq = rcnt[P.ntc];
    P.pos.resize(P.nq);
    P.src.resize(P.nq);
    std::vector<int64_t> fill(rcnt.begin(), rcnt.end() - 1);
    for (int u = 0; u < P.ntop; ++u) {
        const int s0 = P.sn[u], ks = S.sn_k[s0], Ps = P.plen[u];
        for (int r = 0; r < Ps; ++r) {
            const int t = P.ploc[P.poff[u] + r];
            for (int jj = 0; jj < std::min(r + 1, ks); ++jj) {
                const int64_t q = fill[t]++;
                P.pos[q] = P.boff[u] + r + (int64_t)Ps * jj;
                P.src[q] = tloc[S.sn_first[s0] + jj];
            }
        }
    }
    // each row's entries cover one contiguous ascending range of c_T (the
    // postorder above): a segment needs only the column of its first entry
    for (int t = 0; t < P.ntc; ++t) {
        bool ok = rcnt[t + 1] > rcnt[t] && P.src[rcnt[t + 1] - 1] == t;  // ends at its diagonal
        for (int64_t q = rcnt[t]; ok && q < rcnt[t + 1]; ++q) ok = P.src[q] == P.src[rcnt[t]] + (q - rcnt[t]);
        if (!ok) throw std::logic_error("partitioned inverse: a row of L_TT^-1 is not one contiguous column range");
    }
    P.segptr.assign(P.ntc + 1, 0);
    for (int t = 0; t < P.ntc; ++t) {
        for (int64_t q = rcnt[t]; q < rcnt[t + 1]; q += kPiSeg) {
            P.segq.push_back(q);
            P.segc.push_back(P.src[q] - q);
        }
        P.segptr[t + 1] = (int32_t)P.segq.size();
        // a row of one segment is written by it; several segments (or none) leave partial sums
        const bool single = P.segptr[t + 1] - P.segptr[t] == 1;
        if (!single) P.mrows.push_back(t);
        for (int g = P.segptr[t]; g < P.segptr[t + 1]; ++g) P.segdst.push_back(single ? t : ~g);
    }
    P.nseg = (int)P.segq.size();
    P.segq.push_back(P.nq);
    std::vector<int32_t>().swap(P.src);  // carried by segc
    // assembly: the update rows of every subtree root (bottom or middle), per
    // middle column and per top column
    std::vector<std::vector<int64_t>> per(P.ntc), perm(P.mcol.size());
    for (int s = 0; s < nsn; ++s) {
        const int par = S.sn_parent[s];
        if (lev[s] == 2 || par < 0 || lev[par] == lev[s]) continue;  // not a subtree root with a parent
        const int c = S.sn_c[s], k = S.sn_k[s];
        for (int a = 0; a < c - k; ++a) {
            const int row = S.sn_rows[S.sn_rows_off[s] + k + a];
            if (tloc[row] >= 0)
                per[tloc[row]].push_back(S.sn_uoff[s] + a);
            else if (lev[s] == 0 && mloc[row] >= 0)
                perm[mloc[row]].push_back(S.sn_uoff[s] + a);
            else
                throw std::logic_error("partitioned inverse: update row below its subtree's level");
        }
    }
    P.cptr.assign(P.ntc + 1, 0);
    for (int t = 0; t < P.ntc; ++t) {
        P.cpos.insert(P.cpos.end(), per[t].begin(), per[t].end());
        P.cptr[t + 1] = (int32_t)P.cpos.size();
    }
    if (dense) {
        make_sub_plan(S, lev, P);
        P.acptr.assign(P.scols.size() + 1, 0);
        for (size_t c = 0; c < P.scols.size(); ++c) {
            const int ml = mloc[P.scols[c]];
            if (ml >= 0) P.acpos.insert(P.acpos.end(), perm[ml].begin(), perm[ml].end());
            P.acptr[c + 1] = (int32_t)P.acpos.size();
        }
    }
    return P;
}

// ---------------------------------------------------------------------------
// Instance
// ---------------------------------------------------------------------------
struct Instance::Impl {
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // a second stream for the bottom subtrees' solve matrices, which overlap
    // the rest of the factorization (fork after the bottom tasks, join before
    // the top inverse's gather)
    hipStream_t side = nullptr;
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;
    int fork_subm = 1;  // option "fork_sub_m": 0 off, 1 after the middle levels, 2 after the bottom tasks
    int n = 0, nb = 0, m = 0, ns = 0, slack = 0;
    // topology / parameters
    DBuf<int32_t> bus_type, from, to;
    DBuf<double> vset, gs, bs, bc, tap, shift;
    DBuf<uint8_t> status;
    // Ybus
    DBuf<int32_t> yrow, ycol, ycptr, ycon;
    DBuf<double> yG, yB, Pij, Qij;
    DBuf<double> yGt, yBt;  // Ybus values at the transposed positions (right-hand side by location)
    // state & powers
    DBuf<double> V, th, P, Q, Pf, Qf, Pt, Qt;
    // measurements
    DBuf<uint8_t> mcode;
    std::vector<uint8_t> mcode_h;  // host copy (device row order)
    std::vector<uint8_t> status_h;  // host copy of the breaker status (device branch order)
    std::vector<int32_t> br_rows_ptr, br_rows;  // device rows of the measurements on each branch
    std::vector<uint8_t> m_present_h;           // measurement in the structure (device row order)
    std::vector<uint8_t> m_rejected_h;          // set aside as bad data (device row order)
    std::vector<int32_t> row_branch_h;          // device branch of a flow row, -1 for a bus row (device row order)
    // normalized residuals: per device H entry the local row of its column in
    // the row's front, per device row that front; fronts by level (root levels
    // first); the sparse inverse subset on the fronts; row leverages
    struct LnrDev {
        bool ok = false;
        uint64_t plan_epoch = ~0ull;
        DBuf<int32_t> hloc, hsn, parent, levels;
        std::vector<int> lvl_off;
        DBuf<double> Z, lev;
    } lnr;
    bool topo_staged = false;
    DBuf<int32_t> mloc, hrow;
    DBuf<double> z, sigma, rs, hx, r;
    DBuf<double> hs;  // sqrt(w) * Jacobian values in the H pattern (FP64 in both forms)
    // conjugate-gradient refinement (mixed form)
    DBuf<int32_t> hcolp;  // H column indices in permuted state order
    DBuf<double> cg_b, cg_x, cg_r, cg_p, cg_q, cg_z, cg_t, cg_parts;
    DBuf<CgScalars> cg_s;
    DBuf<double> jpart, rpart, mpart;
    DBuf<int> ipart;
    // symbolic structures
    DBuf<int32_t> tptr, tpos, trow, tout, gptr, gloc, gdiag, g_sn_ptr;
    DBuf<FrontDesc> fdesc;
    DBuf<ChildDesc> cdesc;
    DBuf<int32_t> hpos;  // H^T position of each device H entry
    DBuf<double> hr;     // (sqrt(w) H)(sqrt(w) r) products in H^T order
    // device data layout: buses in Cuthill-McKee order from the slack, branches
    // and measurement rows sorted by the layout position of their (from) bus
    std::vector<int32_t> lb, bl, lk, kl, lm, ml;
    std::vector<int32_t> qmap;    // natural H position -> device H position
    std::vector<int32_t> yrow_d;  // device Ybus row pointers
    DBuf<int32_t> ml_d, lm_d;
    DBuf<int32_t> lb_d;             // device position of each bus (insertion order)
    double* state_pin = nullptr;    // pinned [V; theta] in insertion order (2n), read back by get_state
    double* state_pin_dev = nullptr;
    size_t state_pin_bytes = 0;
    int32_t* pin_idx = nullptr;  // pinned staging of indexed telemetry
    double* pin_val = nullptr;
    size_t pin_cap = 0;
    bool pin_busy = false;  // a scatter queued on the stream may still read the staging buffers
    int32_t* pin_idx_dev = nullptr;  // device mappings of the pinned staging
    double* pin_val_dev = nullptr;
    DBuf<double> ztmp;
    DBuf<int2> gpair;
    DBuf<int32_t> sn_first, sn_k, sn_c, child_ptr, child, rmap, rows, perm, pth, pv, lists;
    DBuf<int64_t> foff, rmoff, uoff, rows_off, moff;
    DBuf<DevStatus> st;
    DevStatus* hst = nullptr;      // pinned
    DevStatus* hst_dev = nullptr;  // its device mapping
    // two pinned status slots that k_maxfinish writes in turn (slot seq & 1),
    // with their device mapping: the host reads one iteration's status while
    // the next, already queued, writes the other; published = the last
    // sequence number the host has taken
    DevStatus* hslot = nullptr;
    DevStatus* hslot_dev = nullptr;
    uint32_t published = 0;
    PrecBufs<float> f32;
    PrecBufs<double> f64;
    // per-level launch lists: [level][class] -> (offset into lists, count)
    std::vector<int> lv_off[3], lv_cnt[3];
    int n_levels = 0;
    int level_launches = 0;  // non-empty (level, class) launches of one factor / solve sweep
    // task sweeps: bottom subtrees per workgroup, wide top levels per launch,
    // narrow top tail in one workgroup
    bool tasks_ok = false;
    bool use_tasks = true;  // option: task sweeps when the plan exists
    int force_cut = 0, force_wcut = 0;  // option: factorization plan (0: the cost model's)
    int n_bottom = 0, tail_waves = 0, bottom_waves = kBottomThreads / kWave;
    bool has_tail = false;
    DBuf<int32_t> t_fronts, t_lvl_ptr, t_lvl_off, t_mid;
    std::vector<int> mid_off, mid_cnt;
    int plan_cut = 0, plan_wcut = 0;
    double plan_cost_us = 0, level_cost_us = 0;
    // LDS-resident solve plan (per factor precision: element size differs)
    struct LdsDev {
        bool ok = false;
        DBuf<int32_t> th, meta, lvl, ch, rm, src, mid;
        int n_bottom = 0;
        std::vector<int> tail_tasks, mid_off, mid_cnt;
        std::vector<size_t> tail_bytes;
        size_t bottom_bytes = 0;
        double cost_us = 0;
        int cut = 0, wcut = 0;
    } lds[2];  // [0] float, [1] double
    bool use_lds = true;
    int lds_waves_bottom = 16, lds_waves_tail = 16;
    // partitioned-inverse top solve: bottom subtrees in LDS tasks, L_TT^-1 products on top
    struct PiDev {
        bool ok = false;
        int cut = 0, ntop = 0, ntc = 0, nseg = 0, slice = 0;
        int64_t nblock = 0, nq = 0;
        DBuf<int32_t> sn, plen, ploc, tcol, colsn, segptr, cptr, parent, segdst, mrows, cplen, cjj, ufirst, uk, bunit,
            upath, upoff;
        int nmrows = 0, nunit = 0;
        DBuf<int64_t> boff, poff, pos, segq, segc, cpos, cboff, cpoff;
        int64_t usum = 0, npath = 0;  // sum of subtree root update rows; path rows of the top
        int64_t msub0 = 0, usum0 = 0, ncol0 = 0;  // of the bottom subtrees: entries, update rows, columns
        DBuf<double> part;
        LdsDev bottom[2];  // bottom-subtree LDS tasks, [0] float, [1] double
        // dense bottom subtrees
        bool dense = false;
        int nsub = 0, nsub0 = 0;  // subtrees (bottom ones first)
        DBuf<int32_t> acptr;
        DBuf<int64_t> acpos;
        int64_t msub = 0;
        DBuf<int32_t> sroot, sn_n, sR, scptr, scols, bins, gsrc, gdst, grs;
        DBuf<int64_t> geptr, gbase, gmo;
        DBuf<uint64_t> gmask;
        DBuf<int64_t> smoff, suoff, surow;
        int bin_off[kSubBins + 1] = {0, 0, 0, 0, 0};
        size_t bin_bytes[kSubBins] = {0, 0, 0, 0};
        // per bin, its first bin_early subtrees lie inside the factorization's
        // bottom tasks (split_sub_bins): their matrices can be formed while
        // the rest of the tree factors
        int bin_early[kSubBins] = {0, 0, 0, 0};
        std::vector<int32_t> hbins, hroot;  // host copies: bins, and each subtree's root front
        // subtrees of each level (bottom, middle) by rows R (<= 16, 32, 64, 128)
        // and by columns n (<= 8, 16, 32, 64), for the plan description
        int rhist[2][4] = {}, nhist[2][4] = {};
        // the bottom-subtree products' wave schedule (small subtrees packed)
        DBuf<int32_t> pkind, psub;
        int npw = 0;
        PiSizes sizes() const {
            PiSizes z;
            if (ok) {
                z.nblock = nblock;
                z.nq = nq;
                z.ntc = ntc;
                z.msub = dense ? msub : 0;
            }
            return z;
        }
    } pi;
    bool use_pi = true;
    bool pi_dense = true;   // bottom subtrees as dense solve matrices (else LDS tasks)
    bool pi_mid = true;     // a second level of dense subtrees between the bottom ones and the top
    // middle subtree products: waves per subtree (1: one wave each, k_sub_fwd /
    // k_sub_bwd; 4, 8, 16: one workgroup each; 0: by precision), per direction
    int sub_wg_fwd = 0, sub_wg_bwd = 0;
    int sub_pack = -1;  // bottom-subtree products: small subtrees two or four to a wave (1), not (0), by precision (-1)
    bool meas_by_type = true;  // measurement rows grouped by evaluation class (else by anchor bus only)
    // right-hand side by location (needs the class grouping)
    bool rhs_loc = true;  // asked for
    // mixed form: refactor in FP64, round the solve data to FP32 for the
    // iterations on the cached factor (else an FP32 factorization)
    bool mixed64 = true;
    bool mixed64_off = false;  // this gain's FP64 factor had pivots below 1e-13 G_jj
    DBuf<double> vsave, thsave;  // the state before a mixed-form FP64 refactor
    bool use_loc = false;  // in use (asked for, and the rows are grouped by class)
    DBuf<int32_t> bv_ptr, bi_ptr, br_ptr, ytr, inc_ptr, inc;
    DBuf<double> sbus, sbr;
    int pi_cut = 0;    // fronts per bottom subtree; 0 = default
    int pi_waves = 4;  // waves per bottom-subtree workgroup under the partitioned inverse
    // dependency-driven sweeps (all fronts fit in LDS)
    bool use_flow = false;
    // captured REALTIME iterations (one per precision / weighting variant);
    // retired when launch plans change (plan_epoch) or any device buffer of
    // this instance is (re)allocated (its ledger's epoch)
    struct GraphSlot {
        uint64_t alloc_epoch, plan_epoch;
        int prec, robust, newton, refactor, closing, lead;
        double gamma, vmin, vmax, tol;
        hipGraphExec_t exec;
    };
    std::vector<GraphSlot> graphs;
    uint64_t plan_epoch = 0;
    bool lds_flags = false;  // task solves wait on per-front LDS flags instead of per-level barriers
    bool flow_ok = false;       // every front fits the LDS front of the dataflow kernels
    bool force_levels = false;  // caller asked for level-by-level sweeps
    int max_c = 0;
    uint32_t epoch = 0;
    DBuf<uint32_t> done;
    DBuf<int32_t> order_fwd, order_bwd, sn_parent;
    // the last solve's injections, flows, h(x) and residuals are not evaluated
    // yet (eval_outputs); zsnap holds z as that solve saw it once telemetry
    // has changed since
    double ahead_ratio = 0.05;  // last contraction of consecutive steps on the cached factor
    bool outputs_pending = false;
    bool zsnap_valid = false;
    DBuf<double> zsnap;
    // profiling events per phase
    hipEvent_t pev[PH_COUNT][2] = {};
    bool pev_used[PH_COUNT] = {};
    int pev_launches[PH_COUNT] = {};
    // staged telemetry batches resident in HBM
    struct Staged {
        DBuf<int32_t> idx;
        DBuf<double> val;
        int count = 0;
    };
    std::vector<Staged> staged;
};

// Each bin of the subtree-matrix launches, reordered so that the subtrees
// whose fronts all lie in the factorization's bottom tasks (a front is in them
// when its subtree holds at most plan_cut fronts) come first, each part
// largest first as before; bin_early[q] counts them.  Those matrices depend on
// the bottom task sweep alone.  Rerun whenever either plan changes.
static void split_sub_bins(Instance::Impl& D, const Symbolic& S, hipStream_t s) {
    Instance::Impl::PiDev& PI = D.pi;
    for (int q = 0; q < kSubBins; ++q) PI.bin_early[q] = 0;
    if (!PI.ok || !PI.dense || PI.hbins.empty() || !D.tasks_ok || D.plan_cut <= 0) return;
    const int nsn = S.nsn;
    for (int32_t r : PI.hroot)
        if (r < 0 || r >= nsn) return;  // a plan of another structure (rebuilt next)
    std::vector<int32_t> size(nsn, 1);
    for (int f = 0; f < nsn; ++f)
        if (S.sn_parent[f] >= 0) size[S.sn_parent[f]] += size[f];
    std::vector<int32_t> order(PI.hbins.size());
    for (int q = 0; q < kSubBins; ++q) {
        const int a = PI.bin_off[q], b = PI.bin_off[q + 1];
        int k = a;
        for (int pass = 0; pass < 2; ++pass)
            for (int i = a; i < b; ++i) {
                const int t = PI.hbins[i];
                const bool early = size[PI.hroot[t]] <= D.plan_cut;
                if (early == (pass == 0)) order[k++] = t;
            }
        for (int i = a; i < b; ++i)
            if (size[PI.hroot[order[i]]] <= D.plan_cut) ++PI.bin_early[q];
    }
    PI.hbins = order;
    PI.bins.upload(order.data(), order.size(), s);  // same size: the buffer stays
    HIPCHECK(hipStreamSynchronize(s));
}

// Task-sweep plan of the factorization: bottom subtrees of at most `cut` fronts
// per workgroup, top levels wider than `wcut` fronts one launch each, the rest
// in one workgroup.
static void build_factor_plan(Instance::Impl& D, const Symbolic& S, int cut, int wcut, Ledger& bytes,
                              hipStream_t s) {
    SweepPlan P = make_plan(S, cut, wcut, D.bottom_waves, D.tail_waves, true);
    D.t_fronts.upload(P.fronts, bytes, s);
    D.t_lvl_ptr.upload(P.lvl_ptr, bytes, s);
    D.t_lvl_off.upload(P.lvl_off, bytes, s);
    std::vector<int32_t> mid = P.mid;
    if (mid.empty()) mid.push_back(0);
    D.t_mid.upload(mid, bytes, s);
    HIPCHECK(hipStreamSynchronize(s));
    D.mid_off = P.mid_off;
    D.mid_cnt = P.mid_cnt;
    D.n_bottom = P.n_bottom;
    D.has_tail = P.has_tail;
    D.plan_cut = cut;
    D.plan_wcut = wcut;
    D.plan_cost_us = P.cost_us;
    D.tasks_ok = true;
    split_sub_bins(D, S, s);
}

// Partitioned-inverse plan for the current pi_cut / pi_waves: bottom LDS tasks
// of both element sizes and the top index lists (D.pi.ok = false when the
// fronts or blocks do not fit).
static void build_pi(Instance::Impl& D, const Symbolic& S, Ledger& bytes, hipStream_t s) {
    // partitioned-inverse top solve: bottom LDS tasks of at most pi_cut fronts
    D.pi.ok = false;
    D.pi.npw = 0;
    for (int lv = 0; lv < 2; ++lv)
        for (int q = 0; q < 4; ++q) D.pi.rhist[lv][q] = D.pi.nhist[lv][q] = 0;
    if (S.nsn > 0 && S.max_c <= kFrontB) {
        const int cut = D.pi_cut > 0 ? D.pi_cut : kPiCutDefault;
        const size_t budget = 64 * 1024;
        const bool dense = D.pi_dense;
        LdsPlan LP[2];
        bool fit = true;
        if (!dense) {
            SweepPlan SP = make_plan(S, cut, -1, D.pi_waves, D.lds_waves_tail, true);
            for (int e = 0; e < 2; ++e)
                fit = fit && make_lds_plan(S, SP, e ? sizeof(double) : sizeof(float), budget, D.pi_waves,
                                           D.lds_waves_tail, true, LP[e]);
        }
        PiPlan PP;
        if (fit) PP = make_pi_plan(S, cut, dense, D.pi_mid);
        if (fit && 3 * (size_t)PP.slice * sizeof(double) <= 65536 && PP.nblock < ((int64_t)1 << 40)) {
            auto up = [&](DBuf<int32_t>& b, std::vector<int32_t> v) {
                if (v.empty()) v.push_back(0);
                b.upload(v, bytes, s);
            };
            auto up64 = [&](DBuf<int64_t>& b, std::vector<int64_t> v) {
                if (v.empty()) v.push_back(0);
                b.upload(v, bytes, s);
            };
            for (int e = 0; e < 2 && !dense; ++e) {
                Instance::Impl::LdsDev& L = D.pi.bottom[e];
                up(L.th, LP[e].th);
                up(L.meta, LP[e].meta);
                up(L.lvl, LP[e].lvl);
                up(L.ch, LP[e].ch);
                up(L.rm, LP[e].rm);
                up(L.src, LP[e].src);
                L.n_bottom = LP[e].n_bottom;
                L.bottom_bytes = LP[e].bottom_bytes;
                L.cut = cut;
                L.ok = true;
            }
            Instance::Impl::PiDev& Q = D.pi;
            up(Q.sn, PP.sn);
            up(Q.plen, PP.plen);
            up(Q.ploc, PP.ploc);
            up(Q.tcol, PP.tcol);
            up(Q.colsn, PP.colsn);
            up(Q.segptr, PP.segptr);
            up(Q.segdst, PP.segdst);
            up(Q.ufirst, PP.ufirst);
            up(Q.uk, PP.uk);
            up(Q.bunit, PP.bunit);
            Q.nunit = (int)PP.bunit.size();
            up(Q.upath, PP.upath);
            up(Q.upoff, PP.upoff);
            up(Q.mrows, PP.mrows);
            Q.nmrows = (int)PP.mrows.size();
            Q.npath = (int64_t)PP.ploc.size();
            Q.usum = Q.usum0 = Q.ncol0 = 0;
            for (int t = 0; t < PP.nsub; ++t) {
                Q.usum += PP.sR[t] - PP.sn_n[t];
                if (t < PP.nsub0) {
                    Q.usum0 += PP.sR[t] - PP.sn_n[t];
                    Q.ncol0 += PP.sn_n[t];
                }
            }
            Q.msub0 = PP.nsub0 < PP.nsub ? PP.smoff[PP.nsub0] : PP.msub;
            up(Q.cplen, PP.cplen);
            up(Q.cjj, PP.cjj);
            up64(Q.cboff, PP.cboff);
            up64(Q.cpoff, PP.cpoff);
            up(Q.cptr, PP.cptr);
            up(Q.parent, PP.parent);
            up64(Q.boff, PP.boff);
            up64(Q.poff, PP.poff);
            up64(Q.pos, PP.pos);
            up64(Q.segq, PP.segq);
            up64(Q.segc, PP.segc);
            up64(Q.cpos, PP.cpos);
            Q.part.alloc(std::max(1, PP.nseg), bytes);
            Q.dense = dense;
            if (dense) {
                for (int e = 0; e < 2; ++e) {
                    Instance::Impl::LdsDev& L = D.pi.bottom[e];
                    for (DBuf<int32_t>* b : {&L.th, &L.meta, &L.lvl, &L.ch, &L.rm, &L.src, &L.mid}) b->release(bytes);
                    L.ok = false;
                    L.n_bottom = 0;
                    L.bottom_bytes = 0;
                }
                up(Q.sroot, PP.sroot);
                up(Q.sn_n, PP.sn_n);
                up(Q.sR, PP.sR);
                up(Q.gsrc, PP.gsrc);
                up(Q.gdst, PP.gdst);
                up64(Q.geptr, PP.geptr);
                up64(Q.gbase, PP.gbase);
                up(Q.grs, PP.grs);
                up64(Q.gmo, PP.gmo);
                {
                    std::vector<uint64_t> v = PP.gmask;
                    if (v.empty()) v.push_back(0);
                    Q.gmask.upload(v, bytes, s);
                }
                up(Q.scptr, PP.scptr);
                up(Q.scols, PP.scols);
                up(Q.bins, PP.bins);
                up64(Q.smoff, PP.smoff);
                up64(Q.suoff, PP.suoff);
                up64(Q.surow, PP.surow);
                Q.nsub = PP.nsub;
                Q.nsub0 = PP.nsub0;
                Q.msub = PP.msub;
                for (int t = 0; t < PP.nsub; ++t) {
                    const int lv = t < PP.nsub0 ? 0 : 1;
                    const int R = PP.sR[t], n = PP.sn_n[t];
                    Q.rhist[lv][R <= 16 ? 0 : R <= 32 ? 1 : R <= 64 ? 2 : 3]++;
                    Q.nhist[lv][n <= 8 ? 0 : n <= 16 ? 1 : n <= 32 ? 2 : 3]++;
                }
                up(Q.acptr, PP.acptr);
                up64(Q.acpos, PP.acpos);
                up(Q.pkind, PP.pkind);
                up(Q.psub, PP.psub);
                Q.npw = (int)PP.pkind.size();
                for (int q = 0; q <= kSubBins; ++q) Q.bin_off[q] = PP.bin_off[q];
                for (int q = 0; q < kSubBins; ++q) Q.bin_bytes[q] = PP.bin_bytes[q];
                Q.hbins = PP.bins;
                Q.hroot = PP.sroot;
            }
            HIPCHECK(hipStreamSynchronize(s));
            Q.cut = cut;
            Q.ntop = PP.ntop;
            Q.ntc = PP.ntc;
            Q.nseg = PP.nseg;
            Q.slice = PP.slice;
            Q.nblock = PP.nblock;
            Q.nq = PP.nq;
            Q.ok = true;
        }
    }
    split_sub_bins(D, S, s);
}

Instance::Instance(int device, int precision) : device_(device), precision_(precision) {
    init_device_scratch(device_);
    HIPCHECK(hipSetDevice(device_));
    d_ = new Impl();
    d_->stream = EnginePool::stream();
    HIPCHECK(hipEventCreate(&d_->ev0));
    HIPCHECK(hipEventCreate(&d_->ev1));
    d_->side = EnginePool::stream();
    HIPCHECK(hipEventCreateWithFlags(&d_->fork_ev, hipEventDisableTiming));
    HIPCHECK(hipEventCreateWithFlags(&d_->join_ev, hipEventDisableTiming));
    d_->hst = static_cast<DevStatus*>(EnginePool::pinned(sizeof(DevStatus)));
    HIPCHECK(hipHostGetDevicePointer((void**)&d_->hst_dev, d_->hst, 0));
    d_->hslot = static_cast<DevStatus*>(EnginePool::pinned(2 * sizeof(DevStatus)));
    HIPCHECK(hipHostGetDevicePointer((void**)&d_->hslot_dev, d_->hslot, 0));
    std::memset(d_->hslot, 0, 2 * sizeof(DevStatus));
    d_->st.alloc(1, bytes_);
    // the iteration sequence starts at zero on both sides (outside any capture)
    HIPCHECK(hipMemsetAsync(d_->st.p, 0, sizeof(DevStatus), d_->stream));
    std::memset(d_->hst, 0, sizeof(DevStatus));
    HIPCHECK(hipStreamSynchronize(d_->stream));
    for (int p = 0; p < PH_COUNT; ++p)
        for (int q = 0; q < 2; ++q) HIPCHECK(hipEventCreate(&d_->pev[p][q]));
}

static void drop_graphs(Instance::Impl& D) {
    for (auto& g : D.graphs) (void)hipGraphExecDestroy(g.exec);
    D.graphs.clear();
}

Instance::~Instance() {
    if (!d_) return;
    (void)hipSetDevice(device_);
    if (d_->stream) (void)hipStreamSynchronize(d_->stream);
    Ledger b = bytes_;
    // DBuf destructors do not free; release explicitly
    auto rel = [&](auto& buf) { buf.release(b); };
    rel(d_->bus_type); rel(d_->from); rel(d_->to); rel(d_->vset); rel(d_->gs); rel(d_->bs); rel(d_->bc);
    rel(d_->tap); rel(d_->shift); rel(d_->status); rel(d_->yrow); rel(d_->ycol); rel(d_->ycptr); rel(d_->ycon);
    rel(d_->yG); rel(d_->yB); rel(d_->yGt); rel(d_->yBt); rel(d_->Pij); rel(d_->Qij); rel(d_->V); rel(d_->th); rel(d_->P); rel(d_->Q);
    rel(d_->Pf); rel(d_->Qf); rel(d_->Pt); rel(d_->Qt); rel(d_->mcode); rel(d_->mloc); rel(d_->hrow);
    rel(d_->hs); rel(d_->hcolp); rel(d_->cg_b); rel(d_->cg_x); rel(d_->cg_r); rel(d_->cg_p); rel(d_->cg_q);
    rel(d_->cg_z); rel(d_->cg_t); rel(d_->cg_parts); rel(d_->cg_s); rel(d_->z); rel(d_->sigma); rel(d_->rs); rel(d_->hx); rel(d_->r); rel(d_->zsnap); rel(d_->jpart); rel(d_->rpart); rel(d_->vsave); rel(d_->thsave);
    rel(d_->mpart); rel(d_->ipart); rel(d_->tptr); rel(d_->tpos); rel(d_->trow);
    rel(d_->tout); rel(d_->ml_d); rel(d_->lm_d); rel(d_->lb_d); rel(d_->ztmp);
    if (d_->state_pin) EnginePool::give_back_pinned(d_->state_pin, d_->state_pin_bytes);
    rel(d_->bv_ptr); rel(d_->bi_ptr); rel(d_->br_ptr); rel(d_->ytr); rel(d_->inc_ptr); rel(d_->inc); rel(d_->sbus);
    rel(d_->sbr);
    if (d_->pin_idx) (void)hipHostFree(d_->pin_idx);
    if (d_->pin_val) (void)hipHostFree(d_->pin_val); rel(d_->hpos); rel(d_->hr);
    rel(d_->gptr); rel(d_->gloc); rel(d_->gdiag); rel(d_->g_sn_ptr); rel(d_->fdesc); rel(d_->cdesc); rel(d_->gpair); rel(d_->sn_first);
    rel(d_->sn_k); rel(d_->sn_c); rel(d_->child_ptr); rel(d_->child); rel(d_->rmap); rel(d_->rows);
    rel(d_->perm); rel(d_->pth); rel(d_->pv); rel(d_->lists); rel(d_->foff); rel(d_->rmoff); rel(d_->uoff);
    rel(d_->rows_off); rel(d_->moff); rel(d_->st); rel(d_->done); rel(d_->order_fwd); rel(d_->order_bwd); rel(d_->sn_parent);
    rel(d_->t_fronts); rel(d_->t_lvl_ptr); rel(d_->t_lvl_off); rel(d_->t_mid);
    for (auto& L : d_->lds) {
        rel(L.th); rel(L.meta); rel(L.lvl); rel(L.ch); rel(L.rm); rel(L.src); rel(L.mid);
    }
    {
        auto& P = d_->pi;
        rel(P.sn); rel(P.plen); rel(P.ploc); rel(P.tcol); rel(P.colsn); rel(P.segptr); rel(P.cptr); rel(P.parent);
        rel(P.segdst); rel(P.ufirst); rel(P.uk); rel(P.bunit); rel(P.upath); rel(P.upoff); rel(P.mrows); rel(P.cplen); rel(P.cjj); rel(P.cboff); rel(P.cpoff);
        rel(P.boff); rel(P.poff); rel(P.pos); rel(P.segq); rel(P.segc); rel(P.cpos); rel(P.part);
        rel(P.sroot); rel(P.sn_n); rel(P.sR); rel(P.gsrc); rel(P.gdst); rel(P.geptr); rel(P.gbase); rel(P.grs); rel(P.gmo); rel(P.gmask); rel(P.scptr);
        rel(P.scols); rel(P.bins); rel(P.smoff); rel(P.suoff); rel(P.surow); rel(P.acptr); rel(P.acpos); rel(P.pkind); rel(P.psub);
        for (auto& L : P.bottom) {
            rel(L.th); rel(L.meta); rel(L.lvl); rel(L.ch); rel(L.rm); rel(L.src); rel(L.mid);
        }
    }
    d_->f32.release(b);
    d_->f64.release(b);
    drop_graphs(*d_);
    for (auto& sb : d_->staged) {
        sb.idx.release(b);
        sb.val.release(b);
    }
    for (int p = 0; p < PH_COUNT; ++p)
        for (int q = 0; q < 2; ++q)
            if (d_->pev[p][q]) (void)hipEventDestroy(d_->pev[p][q]);
    if (d_->hst) EnginePool::give_back_pinned(d_->hst, sizeof(DevStatus));
    if (d_->hslot) EnginePool::give_back_pinned(d_->hslot, 2 * sizeof(DevStatus));
    if (d_->ev0) (void)hipEventDestroy(d_->ev0);
    if (d_->ev1) (void)hipEventDestroy(d_->ev1);
    if (d_->fork_ev) (void)hipEventDestroy(d_->fork_ev);
    if (d_->join_ev) (void)hipEventDestroy(d_->join_ev);
    if (d_->side) {
        (void)hipStreamSynchronize(d_->side);
        EnginePool::give_back(d_->side);
    }
    if (d_->stream) {
        (void)hipStreamSynchronize(d_->stream);
        EnginePool::give_back(d_->stream);
    }
    delete d_;
    d_ = nullptr;
}

void Instance::setup(const HostArrays& h, std::shared_ptr<const Symbolic> Sp) {
    ++d_->plan_epoch;  // launch plans change: captured graphs retire
    HIPCHECK(hipSetDevice(device_));
    StreamScope scope_(d_->stream);
    S_ = std::move(Sp);
    const Symbolic& S = *S_;
    Impl& D = *d_;
    hipStream_t s = D.stream;
    D.n = h.n;
    D.nb = h.nb;
    D.m = h.m;
    D.ns = S.ns;
    D.slack = h.slack;
    if (S.max_c > kVecMax) throw std::runtime_error("front larger than the supported maximum (2048)");
    // ---- device data layout ----------------------------------------------
    const int n = h.n, nb = h.nb, m = h.m;
    D.lb.assign(n, -1);
    D.bl.clear();
    {
        // Cuthill-McKee over the Ybus graph from the slack (then other components)
        std::vector<int32_t> q;
        q.reserve(n);
        for (int root = -1; (int)D.bl.size() < n;) {
            int start = root < 0 ? h.slack : -1;
            if (start < 0)
                for (int b = 0; b < n; ++b)
                    if (D.lb[b] < 0) {
                        start = b;
                        break;
                    }
            root = start;
            D.lb[start] = (int32_t)D.bl.size();
            D.bl.push_back(start);
            for (size_t head = D.bl.size() - 1; head < D.bl.size(); ++head) {
                int b = D.bl[head];
                for (int e = S.yrow[b]; e < S.yrow[b + 1]; ++e) {
                    int j = S.ycol[e];
                    if (D.lb[j] < 0) {
                        D.lb[j] = (int32_t)D.bl.size();
                        D.bl.push_back(j);
                    }
                }
            }
        }
    }
    D.kl.resize(nb);
    for (int k = 0; k < nb; ++k) D.kl[k] = k;
    std::stable_sort(D.kl.begin(), D.kl.end(), [&](int a, int b) { return D.lb[h.from[a]] < D.lb[h.from[b]]; });
    D.lk.assign(nb, 0);
    for (int p = 0; p < nb; ++p) D.lk[D.kl[p]] = p;
    auto anchor = [&](int i) { return is_bus_measurement(h.mtype[i]) ? h.mloc[i] : h.from[h.mloc[i]]; };
    // rows by evaluation class (bus values, injections, branch flows) when
    // asked, so that a wavefront runs one code path; by anchor bus inside a class
    auto mclass = [&](int i) {
        if (!D.meas_by_type) return 0;
        const int t = h.mtype[i];
        return (t == 0 || t == 1) ? 0 : ((t == 2 || t == 3 || t == 7 || t == 8) ? 1 : 2);
    };
    // inside the class of branch flows: by branch (device order) and end, so
    // that the rows of one branch end are contiguous
    auto mkey = [&](int i) -> int64_t {
        if (D.meas_by_type && mclass(i) == 2) return (int64_t)D.lk[h.mloc[i]] * 2 + h.mend[i];
        return D.lb[anchor(i)];
    };
    D.ml.resize(m);
    for (int i = 0; i < m; ++i) D.ml[i] = i;
    std::stable_sort(D.ml.begin(), D.ml.end(), [&](int a, int b) {
        const int ca = mclass(a), cb = mclass(b);
        return ca != cb ? ca < cb : mkey(a) < mkey(b);
    });
    D.lm.assign(m, 0);
    for (int p = 0; p < m; ++p) D.lm[D.ml[p]] = p;
    auto bus_perm = [&](const auto& v) {
        std::vector<std::decay_t<decltype(v[0])>> o(n);
        for (int p = 0; p < n; ++p) o[p] = v[D.bl[p]];
        return o;
    };
    auto br_perm = [&](const auto& v) {
        std::vector<std::decay_t<decltype(v[0])>> o(nb);
        for (int p = 0; p < nb; ++p) o[p] = v[D.kl[p]];
        return o;
    };
    D.slack = D.lb[h.slack];
    D.bus_type.upload(bus_perm(h.bus_type), bytes_, s);
    D.vset.upload(bus_perm(h.vset), bytes_, s);
    {
        std::vector<int32_t> f(nb), t(nb);
        for (int p = 0; p < nb; ++p) {
            f[p] = D.lb[h.from[D.kl[p]]];
            t[p] = D.lb[h.to[D.kl[p]]];
        }
        D.from.upload(f, bytes_, s);
        D.to.upload(t, bytes_, s);
        HIPCHECK(hipStreamSynchronize(s));
    }
    D.gs.upload(br_perm(h.gs), bytes_, s);
    D.bs.upload(br_perm(h.bs), bytes_, s);
    D.bc.upload(br_perm(h.bc), bytes_, s);
    D.tap.upload(br_perm(h.tap), bytes_, s);
    D.shift.upload(br_perm(h.shift), bytes_, s);
    D.status_h = br_perm(h.status);
    D.status.upload(D.status_h, bytes_, s);
    HIPCHECK(hipStreamSynchronize(s));
    {
        D.br_rows_ptr.assign(nb + 1, 0);
        for (int i = 0; i < m; ++i)
            if (!is_bus_measurement(h.mtype[i])) D.br_rows_ptr[h.mloc[i] + 1]++;
        for (int k = 0; k < nb; ++k) D.br_rows_ptr[k + 1] += D.br_rows_ptr[k];
        D.br_rows.assign(D.br_rows_ptr[nb], 0);
        std::vector<int32_t> f(D.br_rows_ptr.begin(), D.br_rows_ptr.end() - 1);
        for (int i = 0; i < m; ++i)
            if (!is_bus_measurement(h.mtype[i])) D.br_rows[f[h.mloc[i]]++] = D.lm[i];
        D.m_present_h.assign(m, 1);
        D.m_rejected_h.assign(m, 0);
        D.row_branch_h.assign(m, -1);
        for (int i = 0; i < m; ++i) {
            if (!h.rejected.empty()) D.m_rejected_h[D.lm[i]] = h.rejected[i];
            if (!is_bus_measurement(h.mtype[i])) D.row_branch_h[D.lm[i]] = D.lk[h.mloc[i]];
        }
        if (!h.m_present.empty())
            for (int i = 0; i < m; ++i) D.m_present_h[D.lm[i]] = h.m_present[i];
        D.topo_staged = false;
    }
    {
        // Ybus rows in bus layout order; entries keep their natural order (it
        // fixes the column order of injection rows of H)
        std::vector<int32_t> yr(n + 1, 0), yc, cp(1, 0), cn;
        yc.reserve(S.ycol.size());
        for (int p = 0; p < n; ++p) {
            const int b = D.bl[p];
            for (int e = S.yrow[b]; e < S.yrow[b + 1]; ++e) {
                yc.push_back(D.lb[S.ycol[e]]);
                for (int c = S.ycptr[e]; c < S.ycptr[e + 1]; ++c) cn.push_back(D.lk[S.ycon[c] >> 2] * 4 + (S.ycon[c] & 3));
                cp.push_back((int32_t)cn.size());
            }
            yr[p + 1] = (int32_t)yc.size();
        }
        if (cn.empty()) cn.push_back(0);
        D.yrow_d = yr;
        D.yrow.upload(yr, bytes_, s);
        D.ycol.upload(yc, bytes_, s);
        D.ycptr.upload(cp, bytes_, s);
        D.ycon.upload(cn, bytes_, s);
        // location layout of the right-hand side (rows grouped by class, then
        // by bus / branch end; see k_meas_loc)
        D.use_loc = D.rhs_loc && D.meas_by_type;
        if (D.use_loc) {
            std::vector<int32_t> tr(std::max<size_t>(1, yc.size()), 0);
            for (int p = 0; p < n; ++p)
                for (int e = yr[p]; e < yr[p + 1]; ++e) {
                    const int q = yc[e];
                    int f = -1;
                    for (int e2 = yr[q]; e2 < yr[q + 1]; ++e2)
                        if (yc[e2] == p) f = e2;
                    if (f < 0) throw std::logic_error("Ybus pattern not symmetric");
                    tr[e] = f;
                }
            std::vector<int32_t> ip(n + 1, 0), inc(std::max(1, 2 * nb), 0);
            for (int k = 0; k < nb; ++k) {
                ip[D.lb[h.from[D.kl[k]]] + 1]++;
                ip[D.lb[h.to[D.kl[k]]] + 1]++;
            }
            for (int p = 0; p < n; ++p) ip[p + 1] += ip[p];
            std::vector<int32_t> fill(ip.begin(), ip.end() - 1);
            for (int k = 0; k < nb; ++k) {
                inc[fill[D.lb[h.from[D.kl[k]]]]++] = 2 * k;
                inc[fill[D.lb[h.to[D.kl[k]]]]++] = 2 * k + 1;
            }
            // row ranges per location, in device row order
            std::vector<int32_t> bv(n + 1, 0), bi(n + 1, 0), br(2 * (size_t)nb + 1, 0);
            int n0 = 0, n1 = 0;
            for (int i = 0; i < m; ++i) {
                const int c = mclass(i);
                if (c == 0) {
                    bv[D.lb[h.mloc[i]] + 1]++;
                    ++n0;
                } else if (c == 1) {
                    bi[D.lb[h.mloc[i]] + 1]++;
                    ++n1;
                } else {
                    br[2 * (size_t)D.lk[h.mloc[i]] + h.mend[i] + 1]++;
                }
            }
            bi[0] = n0;
            br[0] = n0 + n1;
            for (int p = 0; p < n; ++p) {
                bv[p + 1] += bv[p];
                bi[p + 1] += bi[p];
            }
            for (size_t q = 0; q < 2 * (size_t)nb; ++q) br[q + 1] += br[q];
            // every row must sit inside its location's range
            for (int r = 0; r < m; ++r) {
                const int i = D.ml[r], c = mclass(i);
                const size_t loc = c == 2 ? 2 * (size_t)D.lk[h.mloc[i]] + h.mend[i] : (size_t)D.lb[h.mloc[i]];
                const std::vector<int32_t>& ptr = c == 0 ? bv : (c == 1 ? bi : br);
                if (r < ptr[loc] || r >= ptr[loc + 1]) throw std::logic_error("measurement rows not grouped by location");
            }
            D.ytr.upload(tr, bytes_, s);
            D.inc_ptr.upload(ip, bytes_, s);
            D.inc.upload(inc, bytes_, s);
            D.bv_ptr.upload(bv, bytes_, s);
            D.bi_ptr.upload(bi, bytes_, s);
            D.br_ptr.upload(br, bytes_, s);
            D.sbus.alloc(4 * (size_t)std::max(1, n), bytes_);
            D.sbr.alloc(4 * (size_t)std::max(1, nb), bytes_);
        }
        HIPCHECK(hipStreamSynchronize(s));
    }
    size_t nY = S.ycol.size();
    if (D.yG.n != nY) {
        D.yG.alloc(nY, bytes_);
        D.yB.alloc(nY, bytes_);
        D.Pij.alloc(nY, bytes_);
        D.Qij.alloc(nY, bytes_);
    }
    if (D.use_loc && D.yGt.n != nY) {
        D.yGt.alloc(nY, bytes_);
        D.yBt.alloc(nY, bytes_);
    }
    if (D.V.n != (size_t)h.n) {
        std::vector<double> V0(h.n, 1.0), t0(h.n, 0.0);
        for (int p = 0; p < h.n; ++p)
            if (h.bus_type[D.bl[p]] != 0) V0[p] = h.vset[D.bl[p]];
        D.V.upload(V0, bytes_, s);
        D.th.upload(t0, bytes_, s);
        D.P.alloc(h.n, bytes_);
        D.Q.alloc(h.n, bytes_);
        HIPCHECK(hipStreamSynchronize(s));
    }
    if (D.Pf.n != (size_t)std::max(1, h.nb)) {
        D.Pf.alloc(std::max(1, h.nb), bytes_);
        D.Qf.alloc(std::max(1, h.nb), bytes_);
        D.Pt.alloc(std::max(1, h.nb), bytes_);
        D.Qt.alloc(std::max(1, h.nb), bytes_);
    }
    {
        std::vector<uint8_t> code(m);
        std::vector<int32_t> loc(m), hr(m + 1, 0);
        std::vector<double> sg(m), zz(m);
        for (int p = 0; p < m; ++p) {
            const int i = D.ml[p];
            const bool bm = is_bus_measurement(h.mtype[i]);
            const int l = h.mloc[i];
            code[p] = meas_code(h.mtype[i], h.mend[i], h.active[i], bm ? 0 : (h.from[l] < h.to[l] ? 1 : 0));
            loc[p] = bm ? D.lb[l] : D.lk[l];
            sg[p] = h.sigma[i];
            zz[p] = h.z[i];
            hr[p + 1] = hr[p] + (S.hrow[i + 1] - S.hrow[i]);
        }
        D.mcode.upload(code, bytes_, s);
        D.mcode_h = code;
        D.mloc.upload(loc, bytes_, s);
        D.sigma.upload(sg, bytes_, s);
        D.z.upload(zz, bytes_, s);
        D.hrow.upload(hr, bytes_, s);
        D.ml_d.upload(D.ml, bytes_, s);
        D.lm_d.upload(D.lm, bytes_, s);
        D.lb_d.upload(D.lb, bytes_, s);
        HIPCHECK(hipStreamSynchronize(s));
        D.qmap.assign(S.hcol.size(), 0);
        for (int i = 0; i < m; ++i)
            for (int q = S.hrow[i]; q < S.hrow[i + 1]; ++q) D.qmap[q] = hr[D.lm[i]] + (q - S.hrow[i]);
    }
    if (D.rs.n != (size_t)std::max(1, h.m)) {
        D.rs.alloc(std::max(1, h.m), bytes_);
        D.hx.alloc(std::max(1, h.m), bytes_);
        D.r.alloc(std::max(1, h.m), bytes_);
        D.zsnap.alloc(std::max(1, h.m), bytes_);
    }
    {
        // J partials: one per block of rows (k_meas) or of buses and branches (k_meas_loc)
        const size_t nparts = std::max(nblk(std::max(1, h.m)), nblk((int64_t)h.n + h.nb));
        if (D.jpart.n != nparts) {
            D.jpart.alloc(nparts, bytes_);
            D.rpart.alloc(nparts, bytes_);
            D.ipart.alloc(nparts, bytes_);
        }
    }
    D.mpart.alloc(nblk(std::max(1, h.n)), bytes_);
    D.hs.alloc(std::max<size_t>(1, S.hcol.size()), bytes_);
    {
        std::vector<int32_t> hp(std::max<size_t>(1, S.hcol.size()), 0);
        for (size_t q = 0; q < S.hcol.size(); ++q) hp[D.qmap[q]] = S.iperm[S.hcol[q]];
        D.hcolp.upload(hp, bytes_, s);
        HIPCHECK(hipStreamSynchronize(s));
        const size_t nsv = std::max(1, S.ns);
        for (DBuf<double>* v : {&D.cg_b, &D.cg_x, &D.cg_r, &D.cg_p, &D.cg_q, &D.cg_z}) v->alloc(nsv, bytes_);
        D.cg_t.alloc(std::max(1, h.m), bytes_);
        D.cg_parts.alloc(nblk(nsv), bytes_);
        D.cg_s.alloc(1, bytes_);
    }
    {
        // H^T lists walked in bus-layout order of the states
        std::vector<int32_t> sorder;
        sorder.reserve(S.ns);
        for (int p = 0; p < n; ++p) {
            const int b = D.bl[p];
            if (S.bus_pth[b] >= 0) sorder.push_back(S.bus_pth[b]);
            sorder.push_back(S.bus_pv[b]);
        }
        std::vector<int32_t> tp(S.ns + 1, 0), tq, tr;
        tq.reserve(S.tpos.size());
        tr.reserve(S.tpos.size());
        std::vector<std::pair<int32_t, int32_t>> ent;
        for (int t = 0; t < S.ns; ++t) {
            const int k = sorder[t];
            ent.clear();
            for (int p = S.tptr[k]; p < S.tptr[k + 1]; ++p) ent.emplace_back(D.lm[S.trow[p]], D.qmap[S.tpos[p]]);
            std::sort(ent.begin(), ent.end());  // by device row: neighbouring rows write neighbouring slots
            for (const auto& e : ent) {
                tq.push_back(e.second);
                tr.push_back(e.first);
            }
            tp[t + 1] = (int32_t)tq.size();
        }
        std::vector<int32_t> hp(std::max<size_t>(1, S.hcol.size()), 0);
        for (size_t p = 0; p < S.hcol.size(); ++p) hp[tq[p]] = (int32_t)p;
        D.hpos.upload(hp, bytes_, s);
        D.hr.alloc(std::max<size_t>(1, S.hcol.size()), bytes_);
        if (tq.empty()) {
            tq.push_back(0);
            tr.push_back(0);
        }
        D.tptr.upload(tp, bytes_, s);
        D.tpos.upload(tq, bytes_, s);
        D.trow.upload(tr, bytes_, s);
        D.tout.upload(sorder, bytes_, s);
        HIPCHECK(hipStreamSynchronize(s));
    }
    D.gptr.upload(S.gpair_ptr, bytes_, s);
    {
        std::vector<int2> gp(S.gpair.size() / 2);
        for (size_t q = 0; q < gp.size(); ++q) gp[q] = make_int2(D.qmap[S.gpair[2 * q]], D.qmap[S.gpair[2 * q + 1]]);
        if (gp.empty()) gp.push_back(make_int2(0, 0));
        D.gpair.upload(gp, bytes_, s);
        HIPCHECK(hipStreamSynchronize(s));
    }
    D.gloc.upload(S.gloc, bytes_, s);
    D.gdiag.upload(S.gdiag, bytes_, s);
    D.g_sn_ptr.upload(S.g_sn_ptr, bytes_, s);
    D.sn_first.upload(S.sn_first, bytes_, s);
    D.sn_k.upload(S.sn_k, bytes_, s);
    D.sn_c.upload(S.sn_c, bytes_, s);
    D.child_ptr.upload(S.sn_child_ptr, bytes_, s);
    {
        std::vector<int32_t> ch = S.sn_child;
        if (ch.empty()) ch.push_back(0);
        D.child.upload(ch, bytes_, s);
        std::vector<int32_t> rm = S.rmap;
        if (rm.empty()) rm.push_back(0);
        D.rmap.upload(rm, bytes_, s);
        HIPCHECK(hipStreamSynchronize(s));
    }
    D.rows.upload(S.sn_rows, bytes_, s);
    D.perm.upload(S.perm, bytes_, s);
    D.pth.upload(bus_perm(S.bus_pth), bytes_, s);
    D.pv.upload(bus_perm(S.bus_pv), bytes_, s);
    D.foff.upload(S.sn_foff, bytes_, s);
    D.moff.upload(S.sn_moff, bytes_, s);
    D.rmoff.upload(S.sn_rmoff, bytes_, s);
    D.uoff.upload(S.sn_uoff, bytes_, s);
    D.rows_off.upload(S.sn_rows_off, bytes_, s);
    {
        std::vector<FrontDesc> fd(std::max<size_t>(1, S.nsn));
        for (int q = 0; q < S.nsn; ++q)
            fd[q] = FrontDesc{S.sn_c[q], S.sn_k[q], S.sn_first[q], S.g_sn_ptr[q], S.g_sn_ptr[q + 1] - S.g_sn_ptr[q],
                              S.sn_child_ptr[q], S.sn_child_ptr[q + 1] - S.sn_child_ptr[q], 0};
        std::vector<ChildDesc> cd(std::max<size_t>(1, S.sn_child.size()));
        for (size_t p = 0; p < S.sn_child.size(); ++p) {
            const int ch = S.sn_child[p];
            cd[p] = ChildDesc{S.sn_c[ch], S.sn_k[ch], 0, 0, S.sn_foff[ch], S.sn_rmoff[ch]};
        }
        D.fdesc.upload(fd, bytes_, s);
        D.cdesc.upload(cd, bytes_, s);
        HIPCHECK(hipStreamSynchronize(s));
    }
    // level lists split by front class
    std::vector<int32_t> lists;
    D.n_levels = S.n_levels;
    for (int c = 0; c < 3; ++c) {
        D.lv_off[c].assign(S.n_levels, 0);
        D.lv_cnt[c].assign(S.n_levels, 0);
    }
    for (int l = 0; l < S.n_levels; ++l)
        for (int cls = 0; cls < 3; ++cls) {
            D.lv_off[cls][l] = (int)lists.size();
            for (int p = S.level_ptr[l]; p < S.level_ptr[l + 1]; ++p) {
                int sn = S.level_sn[p];
                int c = S.sn_c[sn];
                int k = c <= kFrontA ? 0 : (c <= kFrontB ? 1 : 2);
                if (k == cls) lists.push_back(sn);
            }
            D.lv_cnt[cls][l] = (int)lists.size() - D.lv_off[cls][l];
        }
    D.level_launches = 0;
    for (int l = 0; l < S.n_levels; ++l)
        for (int cls = 0; cls < 3; ++cls) D.level_launches += D.lv_cnt[cls][l] ? 1 : 0;
    // task sweeps: pick the cheapest plan, compare with plain level launches
    D.max_c = S.max_c;
    D.tasks_ok = false;
    D.level_cost_us = S.n_levels * (kLaunchUs + kFrontUs);
    if (S.nsn > 0 && S.max_c <= kFrontA) {
        const size_t slice_bytes = (size_t)S.max_c * S.max_c * sizeof(double);
        D.bottom_waves = kBottomThreads / kWave;
        D.tail_waves = (int)std::max<size_t>(1, std::min<size_t>(16, 65536 / std::max<size_t>(1, slice_bytes)));
        int best_cut = 0, best_w = 0;
        double best = D.level_cost_us;
        if (D.force_cut > 0) {  // a caller-chosen plan (option factor_cut / factor_wcut)
            best_cut = D.force_cut;
            best_w = D.force_wcut > 0 ? D.force_wcut : 4;
            best = make_plan(S, best_cut, best_w, D.bottom_waves, D.tail_waves, false).cost_us;
        } else
        // the single-workgroup tail takes levels of at most 16 fronts: wider
        // levels ran there one front per wave in turn, ~10 us a round, while a
        // launch of their own costs one round (measured at 200k buses: tails of
        // levels up to 2-16 fronts wide 2.32 ms per factorization, up to 64 wide
        // 2.54 ms)
        for (int cut : {8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096})
            for (int wc : {2, 4, 8, 16}) {
                SweepPlan P = make_plan(S, cut, wc, D.bottom_waves, D.tail_waves, false);
                if (P.cost_us < best) {
                    best = P.cost_us;
                    best_cut = cut;
                    best_w = wc;
                }
            }
        if (best_cut) {
            build_factor_plan(D, S, best_cut, best_w, bytes_, s);
            D.plan_cost_us = best;
        }
    }
    // LDS-resident solve plans (float and double elements)
    for (int e = 0; e < 2; ++e) {
        Impl::LdsDev& L = D.lds[e];
        L.ok = false;
        if (S.nsn == 0 || S.max_c > kFrontB) continue;
        const size_t elem = e ? sizeof(double) : sizeof(float);
        const size_t budget = 64 * 1024;
        double best = D.level_cost_us;
        int bc = 0, bw = 0;
        for (int cut : {32, 64, 128, 256, 512, 1024})
            for (int wc : {8, 16, 32, 64}) {
                SweepPlan SP = make_plan(S, cut, wc, D.lds_waves_bottom, D.lds_waves_tail, true);
                LdsPlan LP;
                if (make_lds_plan(S, SP, elem, budget, D.lds_waves_bottom, D.lds_waves_tail, false, LP) &&
                    LP.cost_us < best) {
                    best = LP.cost_us;
                    bc = cut;
                    bw = wc;
                }
            }
        if (!bc) continue;
        SweepPlan SP = make_plan(S, bc, bw, D.lds_waves_bottom, D.lds_waves_tail, true);
        LdsPlan LP;
        make_lds_plan(S, SP, elem, budget, D.lds_waves_bottom, D.lds_waves_tail, true, LP);
        auto up = [&](DBuf<int32_t>& b, std::vector<int32_t> v) {
            if (v.empty()) v.push_back(0);
            b.upload(v, bytes_, s);
        };
        up(L.th, LP.th);
        up(L.meta, LP.meta);
        up(L.lvl, LP.lvl);
        up(L.ch, LP.ch);
        up(L.rm, LP.rm);
        up(L.src, LP.src);
        up(L.mid, LP.mid);
        HIPCHECK(hipStreamSynchronize(s));
        L.n_bottom = LP.n_bottom;
        L.tail_tasks = LP.tail_tasks;
        L.tail_bytes = LP.tail_bytes;
        L.bottom_bytes = LP.bottom_bytes;
        L.mid_off = LP.mid_off;
        L.mid_cnt = LP.mid_cnt;
        L.cost_us = best;
        L.cut = bc;
        L.wcut = bw;
        L.ok = true;
    }
    build_pi(D, S, bytes_, s);
    // dependency-driven sweeps: level order lists children before parents
    D.flow_ok = S.nsn > 0 && S.max_c <= kFrontB;
    D.use_flow = D.flow_ok && !D.force_levels;
    if (S.nsn > 0) {
        D.order_fwd.upload(S.level_sn, bytes_, s);
        std::vector<int32_t> rev(S.level_sn.rbegin(), S.level_sn.rend());
        D.order_bwd.upload(rev, bytes_, s);
        D.sn_parent.upload(S.sn_parent, bytes_, s);
        D.done.alloc(S.nsn, bytes_);
        HIPCHECK(hipMemsetAsync(D.done.p, 0, S.nsn * sizeof(uint32_t), s));
        D.epoch = 0;
        HIPCHECK(hipStreamSynchronize(s));
    }
    if (lists.empty()) lists.push_back(0);
    D.lists.upload(lists, bytes_, s);
    HIPCHECK(hipStreamSynchronize(s));
    // precision buffers
    if (precision_ == 1) {
        d_->f32.release(bytes_);
        d_->f64.alloc(S, bytes_, d_->pi.sizes());
    } else {
        d_->f64.release(bytes_);
        d_->f32.alloc(S, bytes_, d_->pi.sizes());
    }
    fp64_factor_active_ = false;
    d_->mixed64_off = false;
    factor_valid_ = false;
    d_->outputs_pending = false;
    d_->zsnap_valid = false;
    if (nY) k_ybus<<<nblk(nY), kBlock, 0, s>>>((int)nY, D.ycptr.p, D.ycon.p, D.gs.p, D.bs.p, D.bc.p, D.tap.p,
                                                D.shift.p, D.status.p, D.yG.p, D.yB.p);
    if (nY && D.use_loc)
        k_ybus_tr<<<nblk(nY), kBlock, 0, s>>>((int)nY, D.ytr.p, D.yG.p, D.yB.p, D.yGt.p, D.yBt.p);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(s));
}

namespace {
template <typename V>
std::vector<typename V::value_type> take(const V& v, const std::vector<int32_t>& idx) {
    std::vector<typename V::value_type> o(idx.size());
    for (size_t p = 0; p < idx.size(); ++p) o[p] = v[idx[p]];
    return o;
}
}  // namespace

namespace {
void eval_outputs(Instance::Impl& D);
void keep_z(Instance::Impl& D);
}  // namespace

void Instance::update_branches(const HostArrays& h) {
    ++d_->plan_epoch;  // launch plans change: captured graphs retire
    HIPCHECK(hipSetDevice(device_));
    StreamScope scope_(d_->stream);
    Impl& D = *d_;
    hipStream_t s = D.stream;
    eval_outputs(D);
    // host sources stay alive until the stream has consumed them
    const std::vector<double> gs = take(h.gs, D.kl), bs = take(h.bs, D.kl), bc = take(h.bc, D.kl),
                              tap = take(h.tap, D.kl), shift = take(h.shift, D.kl);
    D.status_h = take(h.status, D.kl);
    D.gs.upload(gs, bytes_, s);
    D.bs.upload(bs, bytes_, s);
    D.bc.upload(bc, bytes_, s);
    D.tap.upload(tap, bytes_, s);
    D.shift.upload(shift, bytes_, s);
    D.status.upload(D.status_h, bytes_, s);
    // measurement activity (rows of open branches weigh zero)
    for (int p = 0; p < D.m; ++p)
        D.mcode_h[p] = (uint8_t)((D.mcode_h[p] & ~(1u << 5)) | ((h.active[D.ml[p]] ? 1u : 0u) << 5));
    D.mcode.upload(D.mcode_h, bytes_, s);
    HIPCHECK(hipStreamSynchronize(s));
    size_t nY = S_->ycol.size();
    if (nY) k_ybus<<<nblk(nY), kBlock, 0, s>>>((int)nY, D.ycptr.p, D.ycon.p, D.gs.p, D.bs.p, D.bc.p, D.tap.p,
                                                D.shift.p, D.status.p, D.yG.p, D.yB.p);
    if (nY && D.use_loc)
        k_ybus_tr<<<nblk(nY), kBlock, 0, s>>>((int)nY, D.ytr.p, D.yG.p, D.yB.p, D.yGt.p, D.yBt.p);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(s));
    factor_valid_ = false;
}

// One breaker operation: branch k (insertion index) closed or open.  Staged on
// the host; refresh_topology() applies all staged operations at once.
// A breaker operation changes device values only (status, row activity, Ybus):
// launches keep their arguments, so captured iteration graphs stay valid.
void Instance::set_branch_closed(int k, bool closed) {
    Impl& D = *d_;
    if (k < 0 || k >= D.nb) throw std::invalid_argument("branch index out of range");
    D.status_h[D.lk[k]] = closed ? 1 : 0;
    for (int q = D.br_rows_ptr[k]; q < D.br_rows_ptr[k + 1]; ++q) {
        const int p = D.br_rows[q];
        const unsigned on = closed && D.m_present_h[p] && !D.m_rejected_h[p] ? 1u : 0u;
        D.mcode_h[p] = (uint8_t)((D.mcode_h[p] & ~(1u << 5)) | (on << 5));
    }
    D.topo_staged = true;
}

void Instance::set_measurement_rejected(int i, bool rejected) {
    Impl& D = *d_;
    if (i < 0 || i >= D.m) throw std::invalid_argument("measurement index out of range");
    const int p = D.lm[i];
    D.m_rejected_h[p] = rejected ? 1 : 0;
    const int k = D.row_branch_h[p];
    const bool closed = k < 0 || D.status_h[k];
    const unsigned on = closed && D.m_present_h[p] && !rejected ? 1u : 0u;
    D.mcode_h[p] = (uint8_t)((D.mcode_h[p] & ~(1u << 5)) | (on << 5));
    D.topo_staged = true;
}

void Instance::refresh_topology() {
    Impl& D = *d_;
    if (!D.topo_staged) return;
    HIPCHECK(hipSetDevice(device_));
    StreamScope scope_(d_->stream);
    hipStream_t s = D.stream;
    eval_outputs(D);
    D.status.upload(D.status_h, bytes_, s);
    D.mcode.upload(D.mcode_h, bytes_, s);
    const size_t nY = S_->ycol.size();
    if (nY) k_ybus<<<nblk(nY), kBlock, 0, s>>>((int)nY, D.ycptr.p, D.ycon.p, D.gs.p, D.bs.p, D.bc.p, D.tap.p,
                                                D.shift.p, D.status.p, D.yG.p, D.yB.p);
    if (nY && D.use_loc)
        k_ybus_tr<<<nblk(nY), kBlock, 0, s>>>((int)nY, D.ytr.p, D.yG.p, D.yB.p, D.yGt.p, D.yBt.p);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(s));
    D.topo_staged = false;
    factor_valid_ = false;
}

// z arrives in insertion order; the device holds it in row layout order
void Instance::set_z(const double* z) {
    HIPCHECK(hipSetDevice(device_));
    StreamScope scope_(d_->stream);
    Impl& D = *d_;
    if (!D.m) return;
    if (D.ztmp.n != (size_t)D.m) D.ztmp.alloc(D.m, bytes_);
    keep_z(D);
    D.ztmp.upload(z, D.m, D.stream);
    k_gather<<<nblk(D.m), kBlock, 0, D.stream>>>(D.m, D.ml_d.p, D.ztmp.p, D.z.p);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(D.stream));
}

void Instance::set_z_indexed(const int32_t* idx, const double* val, int count) {
    if (count <= 0) return;
    HIPCHECK(hipSetDevice(device_));
    StreamScope scope_(d_->stream);
    Impl& D = *d_;
    // through pinned staging that the scatter kernel reads in place (no copy
    // engine: a large copy of another engine on the device cannot delay it).
    // The previous call's scatter must have read the buffers before they are
    // refilled; the call itself returns without waiting (its scatter runs
    // ahead of the next solve on the same stream).
    if (D.pin_busy) {
        HIPCHECK(hipStreamSynchronize(D.stream));
        D.pin_busy = false;
    }
    if (D.pin_cap < (size_t)count) {
        if (D.pin_idx) HIPCHECK(hipHostFree(D.pin_idx));
        if (D.pin_val) HIPCHECK(hipHostFree(D.pin_val));
        D.pin_idx = nullptr;
        D.pin_val = nullptr;
        HIPCHECK(hipHostMalloc((void**)&D.pin_idx, (size_t)count * sizeof(int32_t), hipHostMallocDefault));
        HIPCHECK(hipHostMalloc((void**)&D.pin_val, (size_t)count * sizeof(double), hipHostMallocDefault));
        HIPCHECK(hipHostGetDevicePointer((void**)&D.pin_idx_dev, D.pin_idx, 0));
        HIPCHECK(hipHostGetDevicePointer((void**)&D.pin_val_dev, D.pin_val, 0));
        D.pin_cap = count;
    }
    std::memcpy(D.pin_idx, idx, (size_t)count * sizeof(int32_t));
    std::memcpy(D.pin_val, val, (size_t)count * sizeof(double));
    keep_z(D);
    k_scatter_z_mapped<<<nblk(count), kBlock, 0, D.stream>>>(count, D.pin_idx_dev, D.pin_val_dev, D.lm_d.p, D.z.p);
    HIPCHECK(hipGetLastError());
    D.pin_busy = true;
}

void Instance::set_state(const double* V, const double* th) {
    HIPCHECK(hipSetDevice(device_));
    StreamScope scope_(d_->stream);
    Impl& D = *d_;
    std::vector<double> v(D.n), t(D.n);
    for (int p = 0; p < D.n; ++p) {
        v[p] = V[D.bl[p]];
        t[p] = th[D.bl[p]];
    }
    eval_outputs(D);
    D.V.upload(v.data(), D.n, D.stream);
    D.th.upload(t.data(), D.n, D.stream);
    HIPCHECK(hipStreamSynchronize(D.stream));
}

namespace {
// device (layout order) -> caller (insertion order)
void fetch_unpermute(const DBuf<double>& src, const std::vector<int32_t>& pos_of, double* dst, hipStream_t s) {
    if (!dst || pos_of.empty()) return;
    std::vector<double> tmp(pos_of.size());
    HIPCHECK(hipMemcpyAsync(tmp.data(), src.p, tmp.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    for (size_t i = 0; i < pos_of.size(); ++i) dst[i] = tmp[pos_of[i]];
}
}  // namespace

void Instance::get_state(double* V, double* th) {
    HIPCHECK(hipSetDevice(device_));
    StreamScope scope_(d_->stream);
    Impl& D = *d_;
    if (!D.n) return;
    const size_t need = 2 * (size_t)D.n * sizeof(double);
    if (D.state_pin_bytes != need) {
        if (D.state_pin) {
            HIPCHECK(hipStreamSynchronize(D.stream));
            EnginePool::give_back_pinned(D.state_pin, D.state_pin_bytes);
        }
        D.state_pin = static_cast<double*>(EnginePool::pinned(need));
        HIPCHECK(hipHostGetDevicePointer((void**)&D.state_pin_dev, D.state_pin, 0));
        D.state_pin_bytes = need;
    }
    k_state_out<<<nblk(D.n), kBlock, 0, D.stream>>>(D.n, D.lb_d.p, D.V.p, D.th.p, D.state_pin_dev);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(D.stream));
    if (V) std::memcpy(V, D.state_pin, (size_t)D.n * sizeof(double));
    if (th) std::memcpy(th, D.state_pin + D.n, (size_t)D.n * sizeof(double));
}

void Instance::flat_start() {
    HIPCHECK(hipSetDevice(device_));
    StreamScope scope_(d_->stream);
    eval_outputs(*d_);
    k_flat<<<nblk(d_->n), kBlock, 0, d_->stream>>>(d_->n, d_->bus_type.p, d_->vset.p, d_->V.p, d_->th.p);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(d_->stream));
}

namespace {

template <typename T>
PrecBufs<T>& bufs(Instance::Impl& D);
template <>
PrecBufs<float>& bufs<float>(Instance::Impl& D) {
    return D.f32;
}
template <>
PrecBufs<double>& bufs<double>(Instance::Impl& D) {
    return D.f64;
}

}  // namespace

// Kernel launch helpers (member-free so they can be templated on T)
namespace {

void launch_power(Instance::Impl& D) {
    k_power<<<nblk(D.n), kBlock, 0, D.stream>>>(D.n, D.yrow.p, D.ycol.p, D.yG.p, D.yB.p, D.V.p, D.th.p, D.P.p,
                                                D.Q.p, D.Pij.p, D.Qij.p, D.st.p);
}

// reset: the pass opens an iteration (clears its status flags); stats: J and
// largest-residual partials per block into jpart / rpart / ipart
void launch_meas_loc(Instance::Impl& D, int robust, double gamma, bool reset = true, bool stats = false) {
    LocArgs A;
    A.n = D.n;
    A.nb = D.nb;
    A.robust = robust;
    A.gamma = gamma;
    A.mcode = D.mcode.p;
    A.z = D.z.p;
    A.sigma = D.sigma.p;
    A.bv_ptr = D.bv_ptr.p;
    A.bi_ptr = D.bi_ptr.p;
    A.br_ptr = D.br_ptr.p;
    A.V = D.V.p;
    A.th = D.th.p;
    A.from = D.from.p;
    A.to = D.to.p;
    A.gs = D.gs.p;
    A.bs = D.bs.p;
    A.bc = D.bc.p;
    A.tap = D.tap.p;
    A.shift = D.shift.p;
    A.sbus = D.sbus.p;
    A.sbr = D.sbr.p;
    A.yrow = D.yrow.p;
    A.ycol = D.ycol.p;
    A.yG = D.yG.p;
    A.yB = D.yB.p;
    A.Pout = D.P.p;
    A.Qout = D.Q.p;
    A.st = D.st.p;
    A.reset = reset ? 1 : 0;
    A.jpart = stats ? D.jpart.p : nullptr;
    A.rpart = stats ? D.rpart.p : nullptr;
    A.ipart = stats ? D.ipart.p : nullptr;
    const int64_t nt = (int64_t)D.n + D.nb;
    if (nt) k_meas_loc<<<nblk(nt), kBlock, 0, D.stream>>>(A);
}

template <typename T>
void launch_rhs_loc(Instance::Impl& D) {
    if (D.n)
        k_rhs_loc<T><<<nblk(D.n), kBlock, 0, D.stream>>>(D.n, D.yrow.p, D.ycol.p, D.yGt.p, D.yBt.p, D.P.p, D.Q.p,
                                                         D.V.p, D.th.p, D.sbus.p, D.sbr.p, D.inc_ptr.p, D.inc.p,
                                                         D.pth.p, D.pv.p, bufs<T>(D).x.p, D.st.p);
}

void launch_meas(Instance::Impl& D, int robust, double gamma, int unit_weights, int want_jac, bool want_out,
                 bool want_rhs = false, bool newton = false, const double* z = nullptr) {
    MeasArgs A;
    A.m = D.m;
    A.n = D.n;
    A.pth = D.pth.p;
    A.robust = robust;
    A.unit_weights = unit_weights;
    A.want_jac = want_jac;
    A.newton = newton ? 1 : 0;
    A.gamma = gamma;
    A.mcode = D.mcode.p;
    A.mloc = D.mloc.p;
    A.z = z ? z : D.z.p;
    A.sigma = D.sigma.p;
    A.hrow = D.hrow.p;
    A.yrow = D.yrow.p;
    A.ycol = D.ycol.p;
    A.yG = D.yG.p;
    A.yB = D.yB.p;
    A.Pij = D.Pij.p;
    A.Qij = D.Qij.p;
    A.P = D.P.p;
    A.Q = D.Q.p;
    A.V = D.V.p;
    A.th = D.th.p;
    A.from = D.from.p;
    A.to = D.to.p;
    A.gs = D.gs.p;
    A.bs = D.bs.p;
    A.bc = D.bc.p;
    A.tap = D.tap.p;
    A.shift = D.shift.p;
    A.rs = D.rs.p;
    A.hr = want_rhs ? D.hr.p : nullptr;
    A.hpos = D.hpos.p;
    A.hout = want_out ? D.hx.p : nullptr;
    A.rout = want_out ? D.r.p : nullptr;
    A.jpart = D.jpart.p;
    A.rpart = D.rpart.p;
    A.ipart = D.ipart.p;
    if (D.m) k_meas<<<nblk(D.m), kBlock, 0, D.stream>>>(A, D.hs.p);
}

// The last solve's outputs at its state and telemetry: injections, branch
// flows, each row's h(x) and residual.  Evaluated on first request, or before
// the state, the topology or the row activity changes.
void eval_outputs(Instance::Impl& D) {
    if (!D.outputs_pending) return;
    launch_power(D);
    if (D.nb)
        k_flows<<<nblk(D.nb), kBlock, 0, D.stream>>>(D.nb, D.from.p, D.to.p, D.gs.p, D.bs.p, D.bc.p, D.tap.p,
                                                     D.shift.p, D.status.p, D.V.p, D.th.p, D.Pf.p, D.Qf.p, D.Pt.p,
                                                     D.Qt.p);
    launch_meas(D, 0, 1.5, 0, 0, true, false, false, D.zsnap_valid ? D.zsnap.p : nullptr);
    HIPCHECK(hipGetLastError());
    D.outputs_pending = false;
    D.zsnap_valid = false;
}

// Before z changes while the last solve's outputs are unevaluated: keep z as
// that solve saw it (one device copy, only for the first change).
void keep_z(Instance::Impl& D) {
    if (!D.outputs_pending || D.zsnap_valid || !D.m) return;
    HIPCHECK(hipMemcpyAsync(D.zsnap.p, D.z.p, (size_t)D.m * sizeof(double), hipMemcpyDeviceToDevice, D.stream));
    D.zsnap_valid = true;
}

template <typename T>
void launch_spgemm(Instance::Impl& D, const Symbolic& S) {
    if (S.nnz_g)
        k_spgemm<T><<<nblk(S.nnz_g), kBlock, 0, D.stream>>>((int)S.nnz_g, D.gptr.p, D.gpair.p, D.hs.p,
                                                            bufs<T>(D).Gval.p);
}

// fork (optional): called with D.fork_ev recorded on the stream after the
// bottom tasks (fork_at 2) or after the middle levels (fork_at 1), once the
// launch that follows that point is queued (so that it is dispatched first).
template <typename T>
void launch_factor_sweeps(Instance::Impl& D, T piv_tol, void (*fork)(Instance::Impl&) = nullptr, int fork_at = 0) {
    PrecBufs<T>& B = bufs<T>(D);
    FactArgs<T> A;
    A.fdesc = D.fdesc.p;
    A.cdesc = D.cdesc.p;
    A.sn_first = D.sn_first.p;
    A.sn_k = D.sn_k.p;
    A.sn_c = D.sn_c.p;
    A.foff = D.foff.p;
    A.child_ptr = D.child_ptr.p;
    A.child = D.child.p;
    A.rmoff = D.rmoff.p;
    A.rmap = D.rmap.p;
    A.g_sn_ptr = D.g_sn_ptr.p;
    A.gloc = D.gloc.p;
    A.gdiag = D.gdiag.p;
    A.Gval = B.Gval.p;
    A.F = B.F.p;
    A.piv_tol = piv_tol;
    A.st = D.st.p;
    if (!D.use_flow && D.use_tasks && D.tasks_ok) {
        SolveArgs<T> none{};
        const int slice = D.max_c * D.max_c;
        if (D.n_bottom)
            k_sweep_tasks<T, 0, kBottomThreads><<<D.n_bottom, kWave * D.bottom_waves, D.bottom_waves * slice * sizeof(T), D.stream>>>(
                A, none, D.t_fronts.p, D.t_lvl_ptr.p, D.t_lvl_off.p, 0, slice);
        if (fork && fork_at == 2) HIPCHECK(hipEventRecord(D.fork_ev, D.stream));
        for (size_t l = 0; l < D.mid_off.size(); ++l) {
            k_factor_lds<T, kFrontA><<<D.mid_cnt[l], kWave, 0, D.stream>>>(A, D.t_mid.p + D.mid_off[l]);
            if (fork && fork_at == 2 && l == 0) fork(D);
        }
        if (fork && fork_at == 1) HIPCHECK(hipEventRecord(D.fork_ev, D.stream));
        if (D.has_tail)
            k_sweep_tasks<T, 0><<<1, kWave * D.tail_waves, D.tail_waves * slice * sizeof(T), D.stream>>>(
                A, none, D.t_fronts.p, D.t_lvl_ptr.p, D.t_lvl_off.p, D.n_bottom, slice);
        if (fork && (fork_at == 1 || D.mid_off.empty())) fork(D);
        return;
    }
    if (D.use_flow) {
        const uint32_t ep = ++D.epoch;
        k_factor_flow<T><<<D.sn_first.n, kWave, D.max_c * D.max_c * sizeof(T), D.stream>>>(A, D.order_fwd.p, D.done.p,
                                                                                         ep, kSpinLimit);
        return;
    }
    for (int l = 0; l < D.n_levels; ++l) {
        if (D.lv_cnt[0][l]) k_factor_lds<T, kFrontA><<<D.lv_cnt[0][l], kWave, 0, D.stream>>>(A, D.lists.p + D.lv_off[0][l]);
        if (D.lv_cnt[1][l]) k_factor_lds<T, kFrontB><<<D.lv_cnt[1][l], kWave, 0, D.stream>>>(A, D.lists.p + D.lv_off[1][l]);
        if (D.lv_cnt[2][l]) k_factor_global<T><<<D.lv_cnt[2][l], kBlock, 0, D.stream>>>(A, D.lists.p + D.lv_off[2][l]);
    }
}

template <typename T>
SolveArgs<T> solve_args(Instance::Impl& D) {
    PrecBufs<T>& B = bufs<T>(D);
    SolveArgs<T> A;
    A.sn_first = D.sn_first.p;
    A.sn_k = D.sn_k.p;
    A.sn_c = D.sn_c.p;
    A.foff = D.foff.p;
    A.rows_off = D.rows_off.p;
    A.rows = D.rows.p;
    A.child_ptr = D.child_ptr.p;
    A.child = D.child.p;
    A.rmoff = D.rmoff.p;
    A.rmap = D.rmap.p;
    A.uoff = D.uoff.p;
    A.F = B.F.p;
    A.x = B.x.p;
    A.ubuf = B.ubuf.p;
    A.M = B.M.p;
    A.moff = D.moff.p;
    return A;
}

PiArgs pi_args(const Instance::Impl& D) {
    const Instance::Impl::PiDev& P = D.pi;
    return PiArgs{P.sn.p,   P.plen.p,  P.boff.p,  P.poff.p,  P.ploc.p, P.parent.p, P.tcol.p, P.colsn.p,
                  P.ntop,   P.ntc,     P.slice,   P.cboff.p, P.cpoff.p, P.cplen.p,  P.cjj.p,    P.ufirst.p,
                  P.uk.p,   P.bunit.p, P.nunit,   P.upath.p, P.upoff.p};
}

bool pi_active(const Instance::Impl& D) { return D.use_pi && D.pi.ok; }

SubArgs sub_args(const Instance::Impl& D) {
    const Instance::Impl::PiDev& P = D.pi;
    return SubArgs{P.nsub,    P.msub,    P.sroot.p, P.sn_n.p,  P.sR.p,    P.geptr.p, P.gbase.p,
                   P.gsrc.p,  P.gdst.p,  P.gmo.p,   P.gmask.p, P.grs.p,   P.scptr.p, P.scols.p,
                   P.smoff.p, P.acptr.p, P.acpos.p, P.suoff.p, P.surow.p};
}

// Packed bottom-subtree products: by option, else by precision (measured at
// 200k buses: FP32 solve data 20.2 -> 19.2 us forward, 20.2 -> 18.7 us
// backward; FP64 slower packed, 26.5 -> 31.9 us backward)
template <typename T>
bool sub_packed(const Instance::Impl& D) {
    if (!D.pi.npw) return false;
    return D.sub_pack >= 0 ? D.sub_pack != 0 : sizeof(T) == sizeof(float);
}

SubPack sub_pack(const Instance::Impl& D) {
    return SubPack{D.pi.npw, D.pi.pkind.p, D.pi.psub.p};
}

bool sub_active(const Instance::Impl& D) { return pi_active(D) && D.pi.dense && D.pi.nsub > 0; }

// A kernel's dynamic LDS above 64 KiB (gfx950: up to 160 KiB per workgroup)
// is allowed once per process and kernel.
void allow_big_lds(const void* kernel, size_t bytes) {
    static std::mutex mu;
    static std::map<const void*, size_t> done;
    std::lock_guard<std::mutex> lk(mu);
    size_t& have = done[kernel];
    if (have >= bytes) return;
    HIPCHECK(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    have = bytes;
}

// L_TT^-1 blocks of the current factor and their row-grouped copy
// Subtree matrices of bins [first, first + count) of every bin q: part 0 the
// early subtrees of each bin, part 1 the rest, part 2 all.
template <typename T>
void launch_sub_m(Instance::Impl& D, int part, hipStream_t st) {
    Instance::Impl::PiDev& PI = D.pi;
    for (int q = 0; q < kSubBins; ++q) {
        const int a = PI.bin_off[q] + (part == 1 ? PI.bin_early[q] : 0);
        const int b = part == 0 ? PI.bin_off[q] + PI.bin_early[q] : PI.bin_off[q + 1];
        if (b <= a) continue;
        if (PI.bin_bytes[q] > 65536) allow_big_lds((const void*)k_sub_m_rows<T>, PI.bin_bytes[q]);
        k_sub_m_rows<T><<<b - a, kBlock, PI.bin_bytes[q], st>>>(sub_args(D), solve_args<T>(D), PI.bins.p + a,
                                                              bufs<T>(D).msub.p);
    }
}

// The early subtree matrices on the second stream, behind D.fork_ev (recorded
// after the bottom task sweep or later); the factorization goes on on the first.
template <typename T>
void fork_sub_m(Instance::Impl& D) {
    HIPCHECK(hipStreamWaitEvent(D.side, D.fork_ev, 0));
    launch_sub_m<T>(D, 0, D.side);
    HIPCHECK(hipEventRecord(D.join_ev, D.side));
}

// Whether a factorization forks the early subtree matrices (fork_sub_m).
bool sub_m_forks(const Instance::Impl& D) {
    if (!D.side || D.use_flow || !D.use_tasks || !D.tasks_ok || !D.n_bottom || !sub_active(D)) return false;
    for (int q = 0; q < kSubBins; ++q)
        if (D.pi.bin_early[q]) return true;
    return false;
}

template <typename T>
void launch_pi_inverse(Instance::Impl& D, bool forked = false) {
    Instance::Impl::PiDev& PI = D.pi;
    PrecBufs<T>& B = bufs<T>(D);
    if (sub_active(D)) launch_sub_m<T>(D, forked ? 1 : 2, D.stream);
    if (forked) HIPCHECK(hipStreamWaitEvent(D.stream, D.join_ev, 0));  // the early ones joined
    if (!pi_active(D) || PI.ntop == 0) return;
    const size_t per_wave = 3 * (size_t)PI.slice * sizeof(T);
    const int wpb = (int)std::max<size_t>(1, std::min<size_t>(4, 65536 / per_wave));
    k_pi_inv<T><<<(PI.ntop + wpb - 1) / wpb, kWave * wpb, wpb * per_wave, D.stream>>>(pi_args(D), solve_args<T>(D),
                                                                                     B.pinv.p);
    k_pi_gather<T><<<(unsigned)((PI.nq + kBlock - 1) / kBlock), kBlock, 0, D.stream>>>(PI.nq, PI.pos.p, B.pinv.p,
                                                                                       B.pinv_rows.p);
}

// Everything the solves need from a factor: per-front solve matrices, the
// dense subtree matrices and the top inverse blocks (after a factorization,
// or after the solve plan changed on the same factor).
__global__ __launch_bounds__(kBlock) void k_round_f32(int64_t n, const double* __restrict__ a, float* __restrict__ b) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
        b[i] = (float)a[i];
}

// The mixed form on the FP64 factor: the FP32 solve data are the FP64 factor's
// panels, subtree solve matrices and top inverse blocks rounded once per
// factorization.  The solves then read half the bytes, and the iteration
// operator is the FP64 one up to one rounding per entry (no FP32
// factorization error, which grows with the gain's condition number).
void round_solve_data(Instance::Impl& D, Ledger& bytes) {
    PrecBufs<double>& A = D.f64;
    PrecBufs<float>& B = D.f32;
    auto round = [&](const DBuf<double>& a, DBuf<float>& b) {
        if (b.n != a.n) b.alloc(a.n, bytes);
        if (a.n)
            k_round_f32<<<(int)std::min<int64_t>(4096, nblk((int64_t)a.n)), kBlock, 0, D.stream>>>((int64_t)a.n, a.p,
                                                                                                    b.p);
    };
    round(A.F, B.F);
    round(A.msub, B.msub);
    round(A.pinv, B.pinv);
    round(A.pinv_rows, B.pinv_rows);
}

template <typename T>
void launch_solve_prep(Instance::Impl& D, bool forked = false) {
    // the LDS task solves (every front fits) read M of every front; with dense
    // bottom subtrees only the top inverse blocks need M (of the top fronts)
    const bool top_only = sub_active(D);
    const int nsn = top_only ? D.pi.ntop : (int)D.sn_first.n;
    if (nsn && D.max_c <= kFrontB)
        k_front_m<T><<<(nsn + kBlock / kWave - 1) / (kBlock / kWave), kBlock, 0, D.stream>>>(
            nsn, solve_args<T>(D), bufs<T>(D).M.p, top_only ? D.pi.sn.p : nullptr);
    launch_pi_inverse<T>(D, forked);
}

template <typename T>
void launch_factor(Instance::Impl& D, T piv_tol) {
    const bool fork = D.fork_subm > 0 && sub_m_forks(D);
    launch_factor_sweeps<T>(D, piv_tol, fork ? fork_sub_m<T> : nullptr, D.fork_subm);
    launch_solve_prep<T>(D, fork);
}

struct Prof;
void prof_begin(Prof* pf, int ph);
void prof_end(Prof* pf, int ph, int launches);

template <typename T>
void launch_fwd_lds(const Instance::Impl& D, int grid, int waves, size_t bytes, const LdsTaskArgs& LA,
                    const SolveArgs<T>& A, int base) {
    if (D.lds_flags)
        k_fwd_lds<T, true><<<grid, kWave * waves, bytes, D.stream>>>(LA, A, base);
    else
        k_fwd_lds<T, false><<<grid, kWave * waves, bytes, D.stream>>>(LA, A, base);
}

template <typename T>
void launch_bwd_lds(const Instance::Impl& D, int grid, int waves, size_t bytes, const LdsTaskArgs& LA,
                    const SolveArgs<T>& A, int base) {
    if (D.lds_flags)
        k_bwd_lds<T, true><<<grid, kWave * waves, bytes, D.stream>>>(LA, A, base);
    else
        k_bwd_lds<T, false><<<grid, kWave * waves, bytes, D.stream>>>(LA, A, base);
}

// Middle-level products, nw waves per subtree (4, 8 or 16)
template <typename T>
void launch_sub_fwd_wg(const Instance::Impl& D, int nw, int nsub, const SubArgs& B, const SolveArgs<T>& A,
                       const T* M, int t0) {
    if (nw == 4)
        k_sub_fwd_wg<T, 4><<<nsub, 4 * kWave, 0, D.stream>>>(B, A, M, t0);
    else if (nw == 8)
        k_sub_fwd_wg<T, 8><<<nsub, 8 * kWave, 0, D.stream>>>(B, A, M, t0);
    else
        k_sub_fwd_wg<T, 16><<<nsub, 16 * kWave, 0, D.stream>>>(B, A, M, t0);
}

template <typename T>
void launch_sub_bwd_wg(const Instance::Impl& D, int nw, int nsub, const SubArgs& B, const SolveArgs<T>& A,
                       const T* M, int t0) {
    if (nw == 4)
        k_sub_bwd_wg<T, 4><<<nsub, 4 * kWave, 0, D.stream>>>(B, A, M, t0);
    else if (nw == 8)
        k_sub_bwd_wg<T, 8><<<nsub, 8 * kWave, 0, D.stream>>>(B, A, M, t0);
    else
        k_sub_bwd_wg<T, 16><<<nsub, 16 * kWave, 0, D.stream>>>(B, A, M, t0);
}

template <typename T>
void launch_trisolve(Instance::Impl& D, Prof* pf = nullptr) {
    SolveArgs<T> A = solve_args<T>(D);
    if (!D.use_flow && pi_active(D)) {
        // bottom subtrees in LDS tasks, the top by the partitioned inverse
        const Instance::Impl::PiDev& PI = D.pi;
        const Instance::Impl::LdsDev& PB = PI.bottom[sizeof(T) == sizeof(double) ? 1 : 0];
        LdsTaskArgs LA{PB.th.p, PB.meta.p, PB.lvl.p, PB.ch.p, PB.rm.p, PB.src.p, D.st.p};
        PrecBufs<T>& B = bufs<T>(D);
        const int nw = kBlock / kWave;
        const bool dense = sub_active(D);
        prof_begin(pf, PH_FWD);
        if (dense) {
            // bottom subtrees, then the middle ones on the assembled right-hand side
            const int n0 = PI.nsub0, n1 = PI.nsub - PI.nsub0;
            if (n0) {
                prof_begin(pf, PH_SUB_FWD);
                if (sub_packed<T>(D))
                    k_sub_fwd_pk<T><<<(PI.npw + nw - 1) / nw, kBlock, 0, D.stream>>>(sub_args(D), sub_pack(D), A,
                                                                                     B.msub.p);
                else
                    k_sub_fwd<T, 0><<<(n0 + nw - 1) / nw, kBlock, 0, D.stream>>>(sub_args(D), A, B.msub.p, 0, n0);
                prof_end(pf, PH_SUB_FWD, 1);
            }
            if (n1) {
                prof_begin(pf, PH_MID_FWD);
                // by precision (measured at 200k buses): four waves per subtree
                const int wg = D.sub_wg_fwd > 0 ? D.sub_wg_fwd : 4;
                if (wg > 1)
                    launch_sub_fwd_wg<T>(D, wg, n1, sub_args(D), A, B.msub.p, n0);
                else
                    k_sub_fwd<T, 1><<<(n1 + nw - 1) / nw, kBlock, 0, D.stream>>>(sub_args(D), A, B.msub.p, n0,
                                                                                 PI.nsub);
                prof_end(pf, PH_MID_FWD, 1);
            }
        } else if (PB.n_bottom) {
            launch_fwd_lds<T>(D, PB.n_bottom, D.pi_waves, PB.bottom_bytes, LA, A, 0);
        }
        if (PI.ntc) {
            k_pi_assemble<T><<<nblk(PI.ntc), kBlock, 0, D.stream>>>(PI.ntc, PI.tcol.p, PI.cptr.p, PI.cpos.p, A.x,
                                                                   A.ubuf, B.ctop.p);
            prof_begin(pf, PH_PI_FWD);
            const int fw = (PI.nseg + kPiSegWave - 1) / kPiSegWave;  // waves of the forward product
            k_pi_fwd<T><<<(fw + nw - 1) / nw, kBlock, 0, D.stream>>>(
                PI.nseg, PI.segq.p, PI.segc.p, PI.segdst.p, B.pinv_rows.p, B.ctop.p, PI.part.p, B.ytop.p);
            prof_end(pf, PH_PI_FWD, 1);
            if (PI.nmrows)
                k_pi_rowsum<T><<<(PI.nmrows + nw - 1) / nw, kBlock, 0, D.stream>>>(PI.nmrows, PI.mrows.p,
                                                                                   PI.segptr.p, PI.part.p, B.ytop.p);
        }
        prof_end(pf, PH_FWD, (dense || PB.n_bottom ? 1 : 0) + (PI.ntc ? 2 + (PI.nmrows ? 1 : 0) : 0));
        prof_begin(pf, PH_BWD);
        if (PI.ntc) {
            prof_begin(pf, PH_PI_BWD);
            if constexpr (sizeof(T) == 8)
                k_pi_bwd_col<T><<<(PI.ntc + nw - 1) / nw, kBlock, 0, D.stream>>>(pi_args(D), A, B.pinv.p, B.ytop.p);
            else
                k_pi_bwd<T><<<(PI.nunit + nw - 1) / nw, kBlock, 0, D.stream>>>(pi_args(D), A, B.pinv.p, B.ytop.p);
            prof_end(pf, PH_PI_BWD, 1);
        }
        if (dense) {
            // middle subtrees (their update rows are top columns), then the bottom ones
            const int n0 = PI.nsub0, n1 = PI.nsub - PI.nsub0;
            const T* Mrow = B.msub.p + PI.msub;
            if (n1) {
                prof_begin(pf, PH_MID_BWD);
                // by precision (measured at 200k buses): eight waves per subtree
                // for FP32 data, one wave for FP64 (whose 64-wide row loads of the
                // split kernel cost more than its shorter chain saves)
                const int wg = D.sub_wg_bwd > 0 ? D.sub_wg_bwd : (sizeof(T) == 4 ? 8 : 1);
                if (wg > 1)
                    launch_sub_bwd_wg<T>(D, wg, n1, sub_args(D), A, Mrow, n0);
                else
                    k_sub_bwd<T, 1><<<(n1 + nw - 1) / nw, kBlock, 0, D.stream>>>(sub_args(D), A, Mrow, n0, PI.nsub);
                prof_end(pf, PH_MID_BWD, 1);
            }
            if (n0) {
                prof_begin(pf, PH_SUB_BWD);
                if (sub_packed<T>(D))
                    k_sub_bwd_pk<T><<<(PI.npw + nw - 1) / nw, kBlock, 0, D.stream>>>(sub_args(D), sub_pack(D), A,
                                                                                     Mrow);
                else
                    k_sub_bwd<T, 0><<<(n0 + nw - 1) / nw, kBlock, 0, D.stream>>>(sub_args(D), A, Mrow, 0, n0);
                prof_end(pf, PH_SUB_BWD, 1);
            }
        }
        else if (PB.n_bottom)
            launch_bwd_lds<T>(D, PB.n_bottom, D.pi_waves, PB.bottom_bytes, LA, A, 0);
        prof_end(pf, PH_BWD, (dense || PB.n_bottom ? 1 : 0) + (PI.ntc ? 1 : 0));
        return;
    }
    Instance::Impl::LdsDev& LD = D.lds[sizeof(T) == sizeof(double) ? 1 : 0];
    if (!D.use_flow && D.use_lds && LD.ok) {
        LdsTaskArgs LA{LD.th.p, LD.meta.p, LD.lvl.p, LD.ch.p, LD.rm.p, LD.src.p, D.st.p};
        const int nmid = (int)LD.mid_off.size();
        const int nlaunch = (LD.n_bottom ? 1 : 0) + nmid + (int)LD.tail_tasks.size();
        prof_begin(pf, PH_FWD);
        if (LD.n_bottom)
            launch_fwd_lds<T>(D, LD.n_bottom, D.lds_waves_bottom, LD.bottom_bytes, LA, A, 0);
        for (int l = 0; l < nmid; ++l)
            k_fwd<T, kFrontB><<<LD.mid_cnt[l], kWave, 0, D.stream>>>(A, LD.mid.p + LD.mid_off[l]);
        for (size_t c = 0; c < LD.tail_tasks.size(); ++c)
            launch_fwd_lds<T>(D, 1, D.lds_waves_tail, LD.tail_bytes[c], LA, A, LD.tail_tasks[c]);
        prof_end(pf, PH_FWD, nlaunch);
        prof_begin(pf, PH_BWD);
        for (int c = (int)LD.tail_tasks.size() - 1; c >= 0; --c)
            launch_bwd_lds<T>(D, 1, D.lds_waves_tail, LD.tail_bytes[c], LA, A, LD.tail_tasks[c]);
        for (int l = nmid - 1; l >= 0; --l)
            k_bwd<T, kFrontB><<<LD.mid_cnt[l], kWave, 0, D.stream>>>(A, LD.mid.p + LD.mid_off[l]);
        if (LD.n_bottom)
            launch_bwd_lds<T>(D, LD.n_bottom, D.lds_waves_bottom, LD.bottom_bytes, LA, A, 0);
        prof_end(pf, PH_BWD, nlaunch);
        return;
    }
    if (!D.use_flow && D.use_tasks && D.tasks_ok) {
        FactArgs<T> none{};
        const int slice = D.max_c;
        const int nmid = (int)D.mid_off.size();
        prof_begin(pf, PH_FWD);
        if (D.n_bottom)
            k_sweep_tasks<T, 1><<<D.n_bottom, kWave * D.bottom_waves, D.bottom_waves * slice * sizeof(T), D.stream>>>(
                none, A, D.t_fronts.p, D.t_lvl_ptr.p, D.t_lvl_off.p, 0, slice);
        for (int l = 0; l < nmid; ++l)
            k_fwd<T, kFrontA><<<D.mid_cnt[l], kWave, 0, D.stream>>>(A, D.t_mid.p + D.mid_off[l]);
        if (D.has_tail)
            k_sweep_tasks<T, 1><<<1, kWave * D.tail_waves, D.tail_waves * slice * sizeof(T), D.stream>>>(
                none, A, D.t_fronts.p, D.t_lvl_ptr.p, D.t_lvl_off.p, D.n_bottom, slice);
        prof_end(pf, PH_FWD, (D.n_bottom ? 1 : 0) + nmid + (D.has_tail ? 1 : 0));
        prof_begin(pf, PH_BWD);
        if (D.has_tail)
            k_sweep_tasks<T, 2><<<1, kWave * D.tail_waves, D.tail_waves * slice * sizeof(T), D.stream>>>(
                none, A, D.t_fronts.p, D.t_lvl_ptr.p, D.t_lvl_off.p, D.n_bottom, slice);
        for (int l = nmid - 1; l >= 0; --l)
            k_bwd<T, kFrontA><<<D.mid_cnt[l], kWave, 0, D.stream>>>(A, D.t_mid.p + D.mid_off[l]);
        if (D.n_bottom)
            k_sweep_tasks<T, 2><<<D.n_bottom, kWave * D.bottom_waves, D.bottom_waves * slice * sizeof(T), D.stream>>>(
                none, A, D.t_fronts.p, D.t_lvl_ptr.p, D.t_lvl_off.p, 0, slice);
        prof_end(pf, PH_BWD, (D.n_bottom ? 1 : 0) + nmid + (D.has_tail ? 1 : 0));
        return;
    }
    if (D.use_flow) {
        const int nsn = (int)D.sn_first.n;
        const size_t sm = D.max_c * sizeof(T);
        prof_begin(pf, PH_FWD);
        const uint32_t e1 = ++D.epoch;
        k_fwd_flow<T><<<nsn, kWave, sm, D.stream>>>(A, D.order_fwd.p, D.done.p, e1, kSpinLimit, D.st.p);
        prof_end(pf, PH_FWD, 1);
        prof_begin(pf, PH_BWD);
        const uint32_t e2 = ++D.epoch;
        k_bwd_flow<T><<<nsn, kWave, sm, D.stream>>>(A, D.order_bwd.p, D.sn_parent.p, D.done.p, e2, kSpinLimit, D.st.p);
        prof_end(pf, PH_BWD, 1);
        return;
    }
    prof_begin(pf, PH_FWD);
    for (int l = 0; l < D.n_levels; ++l) {
        if (D.lv_cnt[0][l]) k_fwd<T, kFrontA><<<D.lv_cnt[0][l], kWave, 0, D.stream>>>(A, D.lists.p + D.lv_off[0][l]);
        if (D.lv_cnt[1][l]) k_fwd<T, kFrontB><<<D.lv_cnt[1][l], kWave, 0, D.stream>>>(A, D.lists.p + D.lv_off[1][l]);
        if (D.lv_cnt[2][l]) k_fwd<T, kVecMax><<<D.lv_cnt[2][l], kWave, 0, D.stream>>>(A, D.lists.p + D.lv_off[2][l]);
    }
    prof_end(pf, PH_FWD, D.level_launches);
    prof_begin(pf, PH_BWD);
    for (int l = D.n_levels - 1; l >= 0; --l) {
        if (D.lv_cnt[0][l]) k_bwd<T, kFrontA><<<D.lv_cnt[0][l], kWave, 0, D.stream>>>(A, D.lists.p + D.lv_off[0][l]);
        if (D.lv_cnt[1][l]) k_bwd<T, kFrontB><<<D.lv_cnt[1][l], kWave, 0, D.stream>>>(A, D.lists.p + D.lv_off[1][l]);
        if (D.lv_cnt[2][l]) k_bwd<T, kVecMax><<<D.lv_cnt[2][l], kWave, 0, D.stream>>>(A, D.lists.p + D.lv_off[2][l]);
    }
    prof_end(pf, PH_BWD, D.level_launches);
}

template <typename T>
void launch_rhs(Instance::Impl& D, hipStream_t st = nullptr) {
    if (D.ns)
        k_rhs_sum<T><<<nblk((int64_t)D.ns * kRhsLanes), kBlock, 0, st ? st : D.stream>>>(D.ns, D.tptr.p, D.tout.p,
                                                                                         D.hr.p, bufs<T>(D).x.p);
}

// dx source: the factor-precision solve vector, or the FP64 CG solution
template <typename T>
void launch_state_update(Instance::Impl& D, double vmin, double vmax, const T* dx, int skip_on_pivot) {
    k_update<T><<<nblk(D.n), kBlock, 0, D.stream>>>(D.n, D.pth.p, D.pv.p, dx, D.V.p, D.th.p, vmin, vmax, D.mpart.p,
                                                    D.st.p, skip_on_pivot);
}

// The iteration's status: max |dx| of the update and, with stats, J and the
// largest residual from the measurement pass at the updated state.  stop_tol:
// see k_maxfinish.
void launch_finish(Instance::Impl& D, bool stats, double stop_tol = -1.0) {
    const int nj = stats ? nblk((int64_t)D.n + D.nb) : 0;
    k_maxfinish<<<1, kFinish, 0, D.stream>>>(nblk(D.n), D.mpart.p, D.st.p, D.hslot_dev, nj, D.jpart.p, D.rpart.p,
                                             D.ipart.p, stop_tol);
}

// dx source: the factor-precision solve vector, or the FP64 CG solution
template <typename T>
void launch_update(Instance::Impl& D, double vmin, double vmax, const T* dx, int skip_on_pivot) {
    launch_state_update<T>(D, vmin, vmax, dx, skip_on_pivot);
    launch_finish(D, false);
}

void launch_dot(Instance::Impl& D, const double* a, const double* b, double* out) {
    int nb = nblk(D.ns);
    k_dot_part<<<nb, kBlock, 0, D.stream>>>(D.ns, a, b, D.cg_parts.p);
    k_dot_finish<<<1, kBlock, 0, D.stream>>>(nb, D.cg_parts.p, out);
}

// z = M^{-1} r with the factor in precision T
template <typename T>
void launch_precond(Instance::Impl& D, const double* r, double* z) {
    int nb = nblk(D.ns);
    k_convert<double, T><<<nb, kBlock, 0, D.stream>>>(D.ns, r, bufs<T>(D).x.p);
    launch_trisolve<T>(D, nullptr);
    k_convert<T, double><<<nb, kBlock, 0, D.stream>>>(D.ns, bufs<T>(D).x.p, z);
}

// PCG: cg_x ~ G^{-1} b with `steps` conjugate-gradient steps (b = H^T W r in FP64)
template <typename T>
void launch_pcg(Instance::Impl& D, int steps) {
    const int ns = D.ns, nb = nblk(ns);
    CgScalars* cs = D.cg_s.p;
    // b = H^T W r from the products of the measurement pass (the gain weights
    // of the CG operator may differ from W under the Huber-Newton form)
    k_rhs_sum<double><<<nblk((int64_t)ns * kRhsLanes), kBlock, 0, D.stream>>>(ns, D.tptr.p, D.tout.p, D.hr.p,
                                                                              D.cg_b.p);
    k_cg_init<<<nb, kBlock, 0, D.stream>>>(ns, D.cg_b.p, D.cg_x.p, D.cg_r.p);
    launch_precond<T>(D, D.cg_r.p, D.cg_z.p);
    launch_dot(D, D.cg_r.p, D.cg_z.p, &cs->rz[0]);
    k_convert<double, double><<<nb, kBlock, 0, D.stream>>>(ns, D.cg_z.p, D.cg_p.p);
    for (int k = 0; k < steps; ++k) {
        if (D.m) k_hmul<<<nblk(D.m), kBlock, 0, D.stream>>>(D.m, D.hrow.p, D.hcolp.p, D.hs.p, D.cg_p.p, D.cg_t.p);
        k_rhs<double><<<nb, kBlock, 0, D.stream>>>(ns, D.tptr.p, D.tpos.p, D.trow.p, D.tout.p, D.hs.p, D.cg_t.p,
                                                   D.cg_q.p);
        launch_dot(D, D.cg_p.p, D.cg_q.p, &cs->pq[k]);
        k_cg_xr<<<nb, kBlock, 0, D.stream>>>(ns, D.cg_x.p, D.cg_r.p, D.cg_p.p, D.cg_q.p, cs, k);
        if (k + 1 == steps) break;
        launch_precond<T>(D, D.cg_r.p, D.cg_z.p);
        launch_dot(D, D.cg_r.p, D.cg_z.p, &cs->rz[k + 1]);
        k_cg_p<<<nb, kBlock, 0, D.stream>>>(ns, D.cg_p.p, D.cg_z.p, cs, k);
    }
}

int factor_launches(const Instance::Impl& D) {
    int inv = pi_active(D) && D.pi.ntop ? 2 : 0;
    if (sub_active(D))
        for (int q = 0; q < kSubBins; ++q) inv += D.pi.bin_off[q + 1] > D.pi.bin_off[q] ? 1 : 0;
    if (D.use_flow) return 1 + inv;
    if (D.use_tasks && D.tasks_ok) return (D.n_bottom ? 1 : 0) + (int)D.mid_off.size() + (D.has_tail ? 1 : 0) + inv;
    return D.level_launches + inv;
}

void read_status(Instance::Impl& D) {
    HIPCHECK(hipMemcpyAsync(D.hst, D.st.p, sizeof(DevStatus), hipMemcpyDeviceToHost, D.stream));
    HIPCHECK(hipStreamSynchronize(D.stream));
}

struct Prof {
    Instance::Impl& D;
    bool on;
    void begin(int ph) {
        if (on) HIPCHECK(hipEventRecord(D.pev[ph][0], D.stream));
    }
    void end(int ph, int launches) {
        if (!on) return;
        HIPCHECK(hipEventRecord(D.pev[ph][1], D.stream));
        D.pev_used[ph] = true;
        D.pev_launches[ph] = launches;
    }
    // call after the stream has been synchronized
    void collect(KernelTimes& kt) {
        if (!on) return;
        for (int p = 0; p < PH_COUNT; ++p) {
            if (!D.pev_used[p]) continue;
            float ms = 0;
            HIPCHECK(hipEventElapsedTime(&ms, D.pev[p][0], D.pev[p][1]));
            kt.ms[p] += ms;
            kt.launches[p] += D.pev_launches[p];
            kt.calls[p] += 1;
            D.pev_used[p] = false;
        }
    }
};

void prof_begin(Prof* pf, int ph) {
    if (pf) pf->begin(ph);
}
void prof_end(Prof* pf, int ph, int launches) {
    if (pf) pf->end(ph, launches);
}

// One Gauss-Newton iteration.  closing (iterations on the cached factor with
// the right-hand side by location): the iteration ends with the measurement
// pass at the state it leaves -- the residuals and right-hand-side sums of the
// next iteration, and J and the largest residual of that state, reduced into
// the status -- so the next iteration starts at its right-hand side, and a
// converged solve needs no separate pass for J.  lead: the sums at the current
// state are not there yet (first such iteration of a solve, or after any
// other kind of iteration): the iteration opens with the pass as well.
template <typename T>
void iteration(Instance::Impl& D, const Symbolic& S, const SolveParams& p, bool refactor, T piv_tol, Prof& pf,
               bool newton, bool closing = false, bool lead = true) {
    // row-major Jacobian only where G or the CG product needs it; the RHS
    // products always (by location when no Jacobian values are needed: the
    // injections are then computed inside k_meas_loc)
    const bool direct = p.pcg_steps <= 0;
    const bool loc = D.use_loc && !refactor && direct;
    closing = closing && loc;
    if (!loc) {
        pf.begin(PH_POWER);
        launch_power(D);
        pf.end(PH_POWER, 1);
    }
    if (!closing || lead) {
        const int ph = closing ? PH_MEAS_LEAD : PH_MEAS;
        pf.begin(ph);
        if (loc)
            launch_meas_loc(D, p.robust ? 1 : 0, p.gamma);
        else
            launch_meas(D, p.robust ? 1 : 0, p.gamma, 0, (refactor || !direct) ? 1 : 0, false, true, newton);
        pf.end(ph, 1);
    }
    // a refactoring iteration whose factorization forks the subtree matrices:
    // its right-hand side goes first on the second stream, beside G and the
    // factor (no kernel of theirs touches x), and joins with the fork
    const bool side_rhs = refactor && direct && !loc && D.fork_subm > 0 && sub_m_forks(D);
    if (side_rhs) {
        HIPCHECK(hipEventRecord(D.fork_ev, D.stream));
        HIPCHECK(hipStreamWaitEvent(D.side, D.fork_ev, 0));
        launch_rhs<T>(D, D.side);
    }
    if (refactor) {
        pf.begin(PH_SPGEMM);
        launch_spgemm<T>(D, S);
        pf.end(PH_SPGEMM, 1);
        pf.begin(PH_FACTOR);
        launch_factor<T>(D, piv_tol);
        pf.end(PH_FACTOR, factor_launches(D));
    }
    if (p.pcg_steps > 0) {
        pf.begin(PH_PCG);
        launch_pcg<T>(D, std::min(p.pcg_steps, kMaxCg));
        pf.end(PH_PCG, 0);
        pf.begin(PH_UPDATE);
        launch_update<double>(D, p.vmin, p.vmax, D.cg_x.p, 0);
        pf.end(PH_UPDATE, 2);
        return;
    }
    pf.begin(PH_RHS);
    if (loc)
        launch_rhs_loc<T>(D);
    else if (!side_rhs)
        launch_rhs<T>(D);
    pf.end(PH_RHS, side_rhs ? 0 : 1);
    launch_trisolve<T>(D, &pf);
    const int skip_on_pivot = sizeof(T) == sizeof(double) ? 1 : 0;
    if (!closing) {
        pf.begin(PH_UPDATE);
        launch_update<T>(D, p.vmin, p.vmax, bufs<T>(D).x.p, skip_on_pivot);
        pf.end(PH_UPDATE, 2);
        return;
    }
    pf.begin(PH_UPDATE);
    launch_state_update<T>(D, p.vmin, p.vmax, bufs<T>(D).x.p, skip_on_pivot);
    pf.end(PH_UPDATE, 1);
    pf.begin(PH_MEAS);
    launch_meas_loc(D, p.robust ? 1 : 0, p.gamma, false, true);
    pf.end(PH_MEAS, 1);
    pf.begin(PH_FINISH);
    launch_finish(D, true, p.tol);
    pf.end(PH_FINISH, 1);
}

// One Gauss-Newton iteration (direct solve) as a HIP graph: the launches of
// iteration() -- on the cached factor, or with the gain product, the
// factorization and its solve data when it refactors (PRECISION, Huber, the
// batched areas) -- are captured once per variant and replayed with one
// launch.  The captured work holds no memset node and no host-varying argument
// (the dependency-driven sweeps, whose epochs change per launch, never come
// here).
template <typename T>
hipGraphExec_t iteration_graph(Instance::Impl& D, const Symbolic& S, const SolveParams& p, T piv_tol, bool newton,
                               bool refactor, uint64_t ae, SolveOutcome& out, bool closing, bool lead,
                               hipGraphExec_t running = nullptr) {
    const uint64_t pe = D.plan_epoch;
    const int prec = (int)sizeof(T);
    hipGraphExec_t exec = nullptr;
    for (size_t k = 0; k < D.graphs.size();) {
        Instance::Impl::GraphSlot& g = D.graphs[k];
        if (g.alloc_epoch != ae || g.plan_epoch != pe) {
            (void)hipGraphExecDestroy(g.exec);
            D.graphs.erase(D.graphs.begin() + k);
            continue;
        }
        if (g.prec == prec && g.robust == p.robust && g.newton == (int)newton && g.refactor == (int)refactor &&
            g.closing == (int)closing && g.lead == (int)lead && g.gamma == p.gamma && g.vmin == p.vmin &&
            g.vmax == p.vmax && g.tol == p.tol)
            exec = g.exec;
        ++k;
    }
    if (!exec) {
        Prof off{D, false};
        hipGraph_t graph = nullptr;
        HIPCHECK(hipStreamBeginCapture(D.stream, hipStreamCaptureModeThreadLocal));
        try {
            iteration<T>(D, S, p, refactor, piv_tol, off, newton, closing, lead);
        } catch (...) {
            (void)hipStreamEndCapture(D.stream, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        HIPCHECK(hipStreamEndCapture(D.stream, &graph));
        const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        HIPCHECK(e);
        if (D.graphs.size() >= 12) {  // the oldest one not queued on the stream right now goes
            const size_t k = D.graphs.front().exec == running ? 1 : 0;
            (void)hipGraphExecDestroy(D.graphs[k].exec);
            D.graphs.erase(D.graphs.begin() + k);
        }
        D.graphs.push_back({ae, pe, prec, p.robust, (int)newton, (int)refactor, (int)closing, (int)lead, p.gamma,
                            p.vmin, p.vmax, p.tol, exec});
        out.graph_captures++;
    }
    return exec;
}

// k_maxfinish, an iteration's last kernel, writes the status into slot
// (seq & 1) of the pinned status block and then the sequence number: the host
// waits for that number instead of a stream synchronization, which wakes it
// several microseconds later (the GPU idles meanwhile unless the next
// iteration is already queued).  A status never seen within the spin budget
// falls back to the synchronization.  The status is copied out of its slot
// at once; the next iteration writes the other slot.
DevStatus take_status(Instance::Impl& D, uint32_t want, bool spin, SolveOutcome& out) {
    DevStatus* slot = D.hslot + (want & 1);
    volatile uint32_t* seq = &slot->seq;
    if (spin) {
        const auto t0 = std::chrono::steady_clock::now();
        for (uint32_t k = 0; *seq != want; ++k) {
            if ((k & 1023) == 1023 && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(200)) {
                HIPCHECK(hipStreamSynchronize(D.stream));
                out.seq_fallbacks++;
                break;
            }
        }
    }
    std::atomic_thread_fence(std::memory_order_acquire);
    DevStatus cur;
    std::memcpy(&cur, slot, sizeof(DevStatus));
    if (cur.seq != want) throw GpuError("iteration status " + std::to_string(want) + " missing (slot holds " +
                                        std::to_string(cur.seq) + ")");
    D.published = want;
    return cur;
}

}  // namespace

SolveOutcome Instance::solve(const SolveParams& p) {
    HIPCHECK(hipSetDevice(device_));
    StreamScope scope_(d_->stream);
    Impl& D = *d_;
    const Symbolic& S = *S_;
    SolveOutcome out;
    Prof pf{D, profiling_};
    auto t0 = std::chrono::steady_clock::now();
    // a new estimate supersedes the previous solve's unevaluated outputs
    D.outputs_pending = false;
    D.zsnap_valid = false;
    if (p.flat_start) k_flat<<<nblk(D.n), kBlock, 0, D.stream>>>(D.n, D.bus_type.p, D.vset.p, D.V.p, D.th.p);
    int retries = 0;
    bool retry = false;
    // contraction of the last two steps of the previous solve on this instance:
    // steps on a cached factor contract more slowly from one iteration to the
    // next (ratios ~0.005, 0.02, 0.04 at 200k buses), so the look-ahead takes
    // the larger of this and the current ratio
    const double tail_ratio = D.ahead_ratio;
    // the last iteration was closed by a measurement pass at the current state:
    // the next one on the cached factor can start at its right-hand side, and
    // its status (last) holds J and the largest residual of this state
    bool closed = false;
    DevStatus last{};
    do {
        if (retry) {
            k_flat<<<nblk(D.n), kBlock, 0, D.stream>>>(D.n, D.bus_type.p, D.vset.p, D.V.p, D.th.p);
            factor_valid_ = false;
            retry = false;
            closed = false;
        }
        bool converged = false, diverged = false, singular = false, timed_out = false;
        int stalls = 0;
        out.iterations = 0;
        out.max_mismatch = 0;
        double last_step = 1e300;
        double prev_step = 0.0;  // the step before last_step, 0 when unknown
        bool newton_off = false;
        bool stale = false;  // REALTIME: the cached factor contracts too slowly, refresh it
        // Iterations on the cached factor that close with the measurement pass
        // are launched one ahead: while iteration it runs, it + 1 is queued
        // behind it unless it is predicted to converge, so that the GPU does
        // not wait for the host between them.  If it does converge, the queued
        // iteration finds the device's stop flag and changes nothing.
        bool ahead = false;   // iteration it was launched by the previous pass
        bool ahead_fp64 = false, ahead_graph = false;
        hipGraphExec_t ahead_exec = nullptr;
        uint32_t ahead_want = 0;
        // closing iterations launched as plain kernels (option queue_plain)
        // are queued and read back like graph replays: by their status's
        // sequence number, with no stream synchronization
        // (mode 1: every such iteration; mode 2: those queued ahead, the
        // solve's first one replaying its graph)
        auto plain_q = [&](bool fp64_, bool closing_, bool ahead_) {
            return (queue_plain_ == 1 || (queue_plain_ == 2 && ahead_)) && closing_ && !profiling_ && !D.use_flow &&
                   (fp64_ || p.pcg_steps <= 0);
        };
        for (int it = 0; it < p.max_iter; ++it) {
            const bool split = precision_ == 0 && !fp64_factor_active_ && D.mixed64 && !D.mixed64_off;
            bool refactor, fp64, newton, graph, closing, lead, queued;
            uint32_t want = 0;  // sequence number of this queued iteration's status
            hipGraphExec_t running = ahead ? ahead_exec : nullptr;
            if (ahead) {
                refactor = false;
                fp64 = ahead_fp64;
                newton = false;
                graph = ahead_graph;
                closing = queued = true;
                lead = false;
                want = ahead_want;
                ahead = false;
            } else {
                refactor = p.robust || !factor_valid_ || p.mode == 1 || stale;
                // mixed form on the FP64 factor: refactor iterations in FP64
                fp64 = precision_ == 1 || fp64_factor_active_ || (split && refactor);
                stale = false;
                // Huber: the IRLS gain until the steps are small, then reduced gain
                // weights of the outliers (same minimiser, faster contraction)
                newton = p.robust == 2 && !newton_off && last_step < kNewtonSwitch;
                // every direct-solve iteration replays a captured graph (one per variant)
                graph = use_graphs_ && !profiling_ && !D.use_flow && (fp64 || p.pcg_steps <= 0);
                // iterations on the cached factor with the right-hand side by
                // location close with the measurement pass at the state they leave
                closing = D.use_loc && !refactor && (fp64 || p.pcg_steps <= 0);
                lead = !closed;
                queued = graph || plain_q(fp64, closing, false);
                if (plain_q(fp64, closing, false)) {
                    graph = false;
                    want = D.published + 1;
                }
                if (fp64) {
                    if (!D.f64.ready) D.f64.alloc(S, bytes_, D.pi.sizes());
                    SolveParams pd = p;
                    pd.pcg_steps = 0;  // the FP64 factor solves the Newton system directly
                    if (split) {  // kept in case this factor turns out singular (below)
                        if (D.vsave.n != D.V.n) D.vsave.alloc(D.V.n, bytes_);
                        if (D.thsave.n != D.th.n) D.thsave.alloc(D.th.n, bytes_);
                        HIPCHECK(hipMemcpyAsync(D.vsave.p, D.V.p, D.V.n * sizeof(double), hipMemcpyDeviceToDevice,
                                                D.stream));
                        HIPCHECK(hipMemcpyAsync(D.thsave.p, D.th.p, D.th.n * sizeof(double),
                                                hipMemcpyDeviceToDevice, D.stream));
                    }
                    if (graph) {
                        hipGraphExec_t g = iteration_graph<double>(D, S, pd, 1e-13, newton, refactor, bytes_.epoch,
                                                                   out, closing, lead);
                        running = g;
                        want = D.published + 1;
                        HIPCHECK(hipGraphLaunch(g, D.stream));
                    } else {
                        iteration<double>(D, S, pd, refactor, 1e-13, pf, newton, closing, lead);
                    }
                    if (split) {
                        if (!D.f32.ready) D.f32.alloc(S, bytes_, D.pi.sizes());
                        round_solve_data(D, bytes_);
                    }
                } else if (graph) {
                    hipGraphExec_t g =
                        iteration_graph<float>(D, S, p, 1e-5f, newton, refactor, bytes_.epoch, out, closing, lead);
                    running = g;
                    want = D.published + 1;
                    HIPCHECK(hipGraphLaunch(g, D.stream));
                } else {
                    iteration<float>(D, S, p, refactor, 1e-5f, pf, newton, closing, lead);
                }
            }
            closed = closing;
            HIPCHECK(hipGetLastError());
            // queue the next iteration behind this one when it would be another
            // closing graph iteration on the cached factor and this one is not
            // predicted to converge (its step from the last known step and the
            // larger of the last contraction and the previous solve's final one,
            // kAheadMargin tol at least)
            if (queued && closing && it + 1 < p.max_iter && p.mode == 0 && !p.robust && factor_valid_ && !stale) {
                const bool nfp64 = precision_ == 1 || fp64_factor_active_;
                const bool nclosing = D.use_loc && (nfp64 || p.pcg_steps <= 0);
                const bool nplain = plain_q(nfp64, nclosing, true);
                const bool ngraph =
                    !nplain && use_graphs_ && !profiling_ && !D.use_flow && (nfp64 || p.pcg_steps <= 0);
                const double ratio = prev_step > 0.0 ? last_step / prev_step : D.ahead_ratio;
                const bool first = it == 0 || last_step >= 1e299;
                const double predicted = first ? 1e300 : last_step * std::min(std::max(ratio, tail_ratio), 1.0);
                bool in_time = true;
                if (p.time_limit_ms > 0)
                    in_time = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                                  .count() < p.time_limit_ms;
                if (nfp64 == fp64 && (ngraph || nplain) && nclosing && in_time && predicted >= kAheadMargin * p.tol) {
                    SolveParams pd = p;
                    if (fp64) pd.pcg_steps = 0;
                    hipGraphExec_t g = nullptr;
                    if (nplain) {
                        if (fp64)
                            iteration<double>(D, S, pd, false, 1e-13, pf, false, true, false);
                        else
                            iteration<float>(D, S, pd, false, 1e-5f, pf, false, true, false);
                    } else {
                        g = fp64 ? iteration_graph<double>(D, S, pd, 1e-13, false, false, bytes_.epoch, out, true,
                                                           false, running)
                                 : iteration_graph<float>(D, S, pd, 1e-5f, false, false, bytes_.epoch, out, true,
                                                          false, running);
                        HIPCHECK(hipGraphLaunch(g, D.stream));
                    }
                    ahead = true;
                    ahead_exec = g;
                    ahead_graph = !nplain;
                    ahead_fp64 = fp64;
                    ahead_want = want + 1;
                    out.ahead++;
                }
            }
            if (!queued) HIPCHECK(hipStreamSynchronize(D.stream));
            // this iteration's status, copied out of its slot (an iteration
            // queued behind writes the other one)
            const DevStatus cur = take_status(D, queued ? want : D.published + 1, queued, out);
            last = cur;
            pf.collect(ktimes_);
            if (cur.piv && newton) {
                // the gain with reduced outlier weights lost rank: IRLS weights from here on
                newton_off = true;
                factor_valid_ = false;
                out.newton_fallbacks++;
                continue;
            }
            if (cur.flow_err) {
                // a dependency wait timed out: redo this iteration level by level
                ++D.plan_epoch;
                D.use_flow = false;
                D.lds_flags = false;
                factor_valid_ = false;
                --it;
                continue;
            }
            if (cur.piv && split) {
                // the FP64 factor of this gain has pivots below 1e-13 G_jj (an
                // unobservable part): back to the state before the step, and the
                // mixed form factors in FP32 from here, substituting such pivots
                HIPCHECK(hipMemcpyAsync(D.V.p, D.vsave.p, D.V.n * sizeof(double), hipMemcpyDeviceToDevice, D.stream));
                HIPCHECK(hipMemcpyAsync(D.th.p, D.thsave.p, D.th.n * sizeof(double), hipMemcpyDeviceToDevice,
                                        D.stream));
                HIPCHECK(hipStreamSynchronize(D.stream));
                closed = false;
                D.mixed64_off = true;
                factor_valid_ = false;
                --it;
                continue;
            }
            if (cur.piv && !fp64) {
                // mixed-form pivots below threshold were substituted (threshold *
                // G_jj): the factor stays SPD; the FP64 residual / RHS of the next
                // iterations (or CG) correct the approximate Newton step
                out.pivots_substituted++;
            } else if (cur.piv) {
                factor_valid_ = false;
                singular = true;
                break;
            }
            if (refactor) {
                factor_valid_ = true;
                out.refactors++;
            }
            out.iterations = it + 1;
            const double prev = out.max_mismatch;
            out.max_mismatch = cur.maxdx;
            // contraction of consecutive steps on the cached factor (the look-ahead's predictor)
            if (!refactor && it > 0 && prev > 0.0) D.ahead_ratio = cur.maxdx / prev;
            prev_step = (!refactor && it > 0) ? prev : 0.0;
            last_step = cur.maxdx;
            // a cached factor (REALTIME) taken at a point the state has since left
            // contracts slowly: refresh it at the current point for the next step
            if (!refactor && it > 0 && cur.maxdx > kRefreshRatio * prev && cur.maxdx >= p.tol) stale = true;
            // the FP32 factor is too coarse for this gain: a factor fresh at this
            // point (or two steps in a row) should contract far better near the
            // solution, so continue with an FP64 factor
            const bool coarse = refactor && !split && it > 0 && cur.maxdx < kLinearRegime &&
                                cur.maxdx > kRefreshRatio * prev && cur.maxdx >= p.tol;
            if (!fp64 && it > 0 && (cur.nan || cur.maxdx > 0.9 * prev || coarse)) {
                if (coarse || ++stalls >= 2 || cur.nan) {
                    fp64_factor_active_ = true;
                    factor_valid_ = false;
                    out.fp64_refactors++;
                    stalls = 0;
                    if (cur.nan) {
                        diverged = true;
                        break;
                    }
                }
            } else {
                stalls = 0;
            }
            if (cur.nan) {
                diverged = true;
                break;
            }
            if (cur.maxdx < p.tol) {
                converged = true;
                break;
            }
            if (p.mode == 0 && p.time_limit_ms > 0 && !ahead) {
                double el = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
                if (el > p.time_limit_ms) {
                    timed_out = true;
                    break;
                }
            }
        }
        // an iteration still queued here follows a converged or diverged one:
        // the device's stop flag keeps it from changing anything
        if (ahead) out.ahead_idle++;
        if (singular)
            out.status = 3;
        else if (diverged)
            out.status = 2;
        else if (converged)
            out.status = 0;
        else
            out.status = 1;
        (void)timed_out;
        if (diverged && retries < p.retries && p.mode != 0) {
            retry = true;
            retries++;
        }
    } while (retry);
    out.retries = retries;
    // J and the largest residual at the final state: from the status of an
    // iteration that a measurement pass closed, else one pass now.  The rest
    // of the outputs (injections, flows, h(x), residuals) are evaluated when
    // first asked for (eval_outputs).
    if (!closed) {
        if (D.use_loc) {
            launch_meas_loc(D, 0, p.gamma, true, true);
            const int nj = nblk((int64_t)D.n + D.nb);
            k_jfinish<<<1, kBlock, 0, D.stream>>>(nj, D.jpart.p, D.rpart.p, D.ipart.p, D.st.p);
        } else {
            launch_power(D);
            launch_meas(D, 0, p.gamma, 0, 0, false);
            if (D.m) k_jfinish<<<1, kBlock, 0, D.stream>>>(nblk(D.m), D.jpart.p, D.rpart.p, D.ipart.p, D.st.p);
        }
        HIPCHECK(hipGetLastError());
        read_status(D);
        last = *D.hst;
    }
    D.outputs_pending = true;
    // z as this solve saw it, copied now: the copy runs while the caller turns
    // to its next request, instead of ahead of that request's telemetry scatter
    keep_z(D);
    out.objective = D.m ? last.J : 0.0;
    out.largest_residual = D.m ? last.rmax : 0.0;
    out.largest_residual_idx = (D.m && last.ridx >= 0) ? D.ml[last.ridx] : -1;
    out.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    out.flags = (D.use_flow ? 1 : 0) | (fp64_factor_active_ ? 2 : 0) |
                (precision_ == 0 && !fp64_factor_active_ && D.mixed64 ? 4 : 0);
    return out;
}

// ---------------------------------------------------------------------------
// Normalized residuals: sparse inverse subset of the gain on the fronts
// ---------------------------------------------------------------------------
// For a supernode with own columns J (k) and below-rows I (u = c - k), the
// front's solve matrix M = [L_JJ^-1; -W] (W = L_IJ L_JJ^-1) and the inverse
// Z = G^-1 restricted to the front's rows:
//   Z_II  from the parent's front (I lies in the parent's rows: rmap)
//   Z_IJ = Z_II (-W)
//   Z_JJ = L_JJ^-T L_JJ^-1 + (-W)^T Z_IJ
// (the block form of the Takahashi recurrences).  Fronts are processed from
// the roots down, one workgroup per front; Z is stored like the fronts
// (c x c, column-major, both triangles).
struct SelInvArgs {
    const int32_t* sn_k;
    const int32_t* sn_c;
    const int64_t* foff;
    const int64_t* moff;
    const int64_t* rmoff;
    const int32_t* rmap;
    const int32_t* parent;
    const double* M;
    double* Z;
};

__global__ __launch_bounds__(kBlock) void k_selinv(SelInvArgs A, const int32_t* __restrict__ list) {
    const int s = list[blockIdx.x];
    const int c = A.sn_c[s], k = A.sn_k[s], u = c - k;
    double* Zs = A.Z + A.foff[s];
    const double* Ms = A.M + A.moff[s];
    const int tid = threadIdx.x;
    if (u > 0) {
        const int p = A.parent[s];
        const int cp = A.sn_c[p];
        const double* Zp = A.Z + A.foff[p];
        const int32_t* rm = A.rmap + A.rmoff[s];
        for (int e = tid; e < u * u; e += kBlock) {
            const int a = e % u, b = e / u;
            Zs[(int64_t)(k + b) * c + (k + a)] = Zp[(int64_t)rm[b] * cp + rm[a]];
        }
        __syncthreads();
        for (int e = tid; e < u * k; e += kBlock) {
            const int a = e % u, j = e / u;
            double acc = 0.0;
            for (int b = 0; b < u; ++b) acc += Zs[(int64_t)(k + b) * c + (k + a)] * Ms[(int64_t)j * c + (k + b)];
            Zs[(int64_t)j * c + (k + a)] = acc;  // row k+a, column j
        }
        __syncthreads();
        for (int e = tid; e < u * k; e += kBlock) {
            const int a = e % u, j = e / u;
            Zs[(int64_t)(k + a) * c + j] = Zs[(int64_t)j * c + (k + a)];
        }
    }
    // Z_JJ reads Z_IJ (columns j < k, rows >= k) and writes rows and columns < k
    for (int e = tid; e < k * k; e += kBlock) {
        const int i = e % k, j = e / k;
        if (i < j) continue;
        double acc = 0.0;
        for (int r = i; r < k; ++r) acc += Ms[(int64_t)i * c + r] * Ms[(int64_t)j * c + r];
        for (int b = 0; b < u; ++b) acc += Ms[(int64_t)i * c + (k + b)] * Zs[(int64_t)j * c + (k + b)];
        Zs[(int64_t)j * c + i] = acc;
        Zs[(int64_t)i * c + j] = acc;
    }
}

// Leverage of each measurement row: h_i Z h_i^T over the front holding the
// row's columns (hs = sqrt(w) H with the base weights, so this is
// h_i G^-1 h_i^T / sigma_i^2).  One thread per row.
__global__ __launch_bounds__(kBlock) void k_leverage(int m, const int32_t* __restrict__ hrow,
                                                     const double* __restrict__ hs, const int32_t* __restrict__ hloc,
                                                     const int32_t* __restrict__ hsn, const int32_t* __restrict__ sn_c,
                                                     const int64_t* __restrict__ foff, const double* __restrict__ Z,
                                                     double* __restrict__ lev) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= m) return;
    const int s = hsn[p];
    if (s < 0) {
        lev[p] = 0.0;
        return;
    }
    const int c = sn_c[s];
    const double* Zs = Z + foff[s];
    const int q0 = hrow[p], q1 = hrow[p + 1];
    double acc = 0.0;
    for (int a = q0; a < q1; ++a) {
        const int64_t col = (int64_t)hloc[a] * c;
        double t = 0.0;
        for (int b = q0; b < q1; ++b) t += hs[b] * Zs[col + hloc[b]];
        acc += hs[a] * t;
    }
    lev[p] = acc;
}

// ---------------------------------------------------------------------------
// Allocation-free solve
// ---------------------------------------------------------------------------
// Scratch of the allocation-free solve: the iteration status record and the
// per-block partials of the reductions (max |dx| per block of buses; objective,
// largest scaled residual and its row per block of measurement rows).  Each
// device holds its own copy until the process ends.  In host code `scratch`
// names no device address: kernels get the address that hipGetSymbolAddress
// returned for the current device (scratch_address).
constexpr size_t kScratchBytes = size_t(1) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
std::mutex g_scratch_mu;
std::map<int, unsigned char*>& scratch_table() {
    static auto* t = new std::map<int, unsigned char*>();  // lives to process end
    return *t;
}

unsigned char* scratch_address(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = scratch_table().find(device);
    return it == scratch_table().end() ? nullptr : it->second;
}

// Offsets into the scratch array (each 256-byte aligned).
struct ScratchLayout {
    size_t st = 0, mpart = 0, jpart = 0, rpart = 0, ipart = 0, total = 0;
};

ScratchLayout scratch_layout(int n, int m, int nb) {
    auto up = [](size_t x) { return (x + 255) / 256 * 256; };
    ScratchLayout L;
    size_t off = 0;
    L.st = off;
    off += up(sizeof(DevStatus));
    L.mpart = off;
    off += up(sizeof(double) * (size_t)nblk(std::max(1, n)));  // k_update: one partial per block of buses
    // J partials: one per block of rows (k_meas) or of buses and branches (k_meas_loc)
    const size_t mp = (size_t)std::max(nblk(std::max(1, m)), nblk((int64_t)n + nb));
    L.jpart = off;
    off += up(sizeof(double) * mp);
    L.rpart = off;
    off += up(sizeof(double) * mp);
    L.ipart = off;
    off += up(sizeof(int) * mp);
    L.total = off;
    return L;
}
}  // namespace

void init_device_scratch(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (scratch_table().count(device)) return;
    HIPCHECK(hipSetDevice(device));
    void* p = nullptr;
    HIPCHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    scratch_table()[device] = static_cast<unsigned char*>(p);
}

size_t device_scratch_bytes() { return kScratchBytes; }

namespace {
// Host plan of the normalized residuals for the current symbolic analysis.
void build_lnr_plan(Instance::Impl& D, const Symbolic& S, Ledger& bytes, hipStream_t s) {
    auto& L = D.lnr;
    const int m = S.m;
    // supernode of each permuted column
    std::vector<int32_t> col_sn(S.ns);
    for (int t = 0; t < S.nsn; ++t)
        for (int j = S.sn_first[t]; j < S.sn_first[t] + S.sn_k[t]; ++j) col_sn[j] = t;
    std::vector<int32_t> hloc(S.hcol.size()), hsn(std::max(1, m), -1);
    for (int i = 0; i < m; ++i) {
        const int q0 = S.hrow[i], q1 = S.hrow[i + 1];
        if (q0 == q1) continue;
        int j0 = S.ns;
        for (int q = q0; q < q1; ++q) j0 = std::min(j0, S.iperm[S.hcol[q]]);
        // the row's columns are a clique of the filled graph: all of them lie
        // in the front of the supernode of the first eliminated one
        const int t = col_sn[j0];
        const int first = S.sn_first[t], k = S.sn_k[t], c = S.sn_c[t];
        const int32_t* rows = S.sn_rows.data() + S.sn_rows_off[t];
        for (int q = q0; q < q1; ++q) {
            const int j = S.iperm[S.hcol[q]];
            int loc;
            if (j >= first && j < first + k) {
                loc = j - first;
            } else {
                const int32_t* e = std::lower_bound(rows + k, rows + c, j);
                if (e == rows + c || *e != j) throw std::logic_error("measurement row outside the front of its first column");
                loc = (int)(e - rows);
            }
            hloc[D.qmap[q]] = loc;
        }
        hsn[D.lm[i]] = t;
    }
    // fronts by level, root levels first (a parent's level is above its children's)
    std::vector<int32_t> lv;
    L.lvl_off.assign(1, 0);
    for (int l = S.n_levels - 1; l >= 0; --l) {
        for (int q = S.level_ptr[l]; q < S.level_ptr[l + 1]; ++q) lv.push_back(S.level_sn[q]);
        L.lvl_off.push_back((int)lv.size());
    }
    if (lv.empty()) lv.push_back(0);
    std::vector<int32_t> par(S.sn_parent);
    if (par.empty()) par.push_back(-1);
    L.hloc.upload(hloc.empty() ? std::vector<int32_t>(1, 0) : hloc, bytes, s);
    L.hsn.upload(hsn, bytes, s);
    L.parent.upload(par, bytes, s);
    L.levels.upload(lv, bytes, s);
    if (L.Z.n != (size_t)std::max<int64_t>(1, S.ftotal)) L.Z.alloc(std::max<int64_t>(1, S.ftotal), bytes);
    if (L.lev.n != (size_t)std::max(1, m)) L.lev.alloc(std::max(1, m), bytes);
    HIPCHECK(hipStreamSynchronize(s));
    L.plan_epoch = D.plan_epoch;
    L.ok = true;
}
}  // namespace

bool Instance::leverages(double* lev, double* r) {
    HIPCHECK(hipSetDevice(device_));
    StreamScope scope_(d_->stream);
    Impl& D = *d_;
    const Symbolic& S = *S_;
    hipStream_t s = D.stream;
    if (!D.lnr.ok || D.lnr.plan_epoch != D.plan_epoch) build_lnr_plan(D, S, bytes_, s);
    if (!D.f64.ready) D.f64.alloc(S, bytes_, D.pi.sizes());
    // the WLS gain at the current state (base weights), factored in FP64
    launch_power(D);
    launch_meas(D, 0, 1.5, 0, 1, true);
    launch_spgemm<double>(D, S);
    launch_factor<double>(D, 1e-13);
    HIPCHECK(hipGetLastError());
    read_status(D);
    if (D.hst->piv) {
        factor_valid_ = false;
        return false;
    }
    // the factor is a valid cached factor at this state; the mixed form's FP32
    // solve data follow it (an FP32-factorization mixed form refactors instead)
    if (precision_ == 1 || fp64_factor_active_) {
        factor_valid_ = true;
    } else if (D.mixed64 && !D.mixed64_off) {
        if (!D.f32.ready) D.f32.alloc(S, bytes_, D.pi.sizes());
        round_solve_data(D, bytes_);
        factor_valid_ = true;
    } else {
        factor_valid_ = false;
    }
    // M = [L_JJ^-1; -L_IJ L_JJ^-1] of every front (the top fronts' are the solve's own)
    const int nsn = (int)S.nsn;
    if (nsn)
        k_front_m<double><<<(nsn + kBlock / kWave - 1) / (kBlock / kWave), kBlock, 0, s>>>(nsn, solve_args<double>(D),
                                                                                          D.f64.M.p, nullptr);
    SelInvArgs A{D.sn_k.p, D.sn_c.p, D.foff.p, D.moff.p, D.rmoff.p, D.rmap.p, D.lnr.parent.p, D.f64.M.p, D.lnr.Z.p};
    for (size_t l = 0; l + 1 < D.lnr.lvl_off.size(); ++l) {
        const int cnt = D.lnr.lvl_off[l + 1] - D.lnr.lvl_off[l];
        if (cnt) k_selinv<<<cnt, kBlock, 0, s>>>(A, D.lnr.levels.p + D.lnr.lvl_off[l]);
    }
    if (D.m)
        k_leverage<<<nblk(D.m), kBlock, 0, s>>>(D.m, D.hrow.p, D.hs.p, D.lnr.hloc.p, D.lnr.hsn.p, D.sn_c.p, D.foff.p,
                                               D.lnr.Z.p, D.lnr.lev.p);
    HIPCHECK(hipGetLastError());
    std::vector<double> lv(D.m), rr(D.m);
    if (D.m) {
        HIPCHECK(hipMemcpyAsync(lv.data(), D.lnr.lev.p, D.m * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHECK(hipMemcpyAsync(rr.data(), D.r.p, D.m * sizeof(double), hipMemcpyDeviceToHost, s));
    }
    HIPCHECK(hipStreamSynchronize(s));
    for (int p = 0; p < D.m; ++p) {
        if (lev) lev[D.ml[p]] = lv[p];
        if (r) r[D.ml[p]] = rr[p];
    }
    return true;
}

SolveOutcome Instance::solve_no_alloc(const SolveParams& p) {
    if (p.mode != 0 || p.robust != 0 || p.pcg_steps > 0)
        throw NoAllocError("the allocation-free solve runs REALTIME WLS iterations on the cached factor "
                           "(no PRECISION mode, no Huber, no CG refinement)");
    if (!factor_valid_) throw NoAllocError("no cached gain factor: solve once with sle_ext_Solve first");
    Impl& D = *d_;
    const ScratchLayout L = scratch_layout(D.n, D.m, D.nb);
    if (L.total > kScratchBytes)
        throw NoAllocError("the model's reduction partials need " + std::to_string(L.total) +
                           " bytes of scratch; the allocation-free solve has " + std::to_string(kScratchBytes));
    unsigned char* base = scratch_address(device_);
    if (!base) throw NoAllocError("the device's scratch address was not looked up (engine not initialized)");
    HIPCHECK(hipSetDevice(device_));
    // the kernels take the scratch addresses where the instance's own buffers
    // go otherwise; the status record (iteration sequence included) is carried
    // over in both directions so that later replays of the instance agree
    DevStatus* own_st = D.st.p;
    double* own_m = D.mpart.p;
    double* own_j = D.jpart.p;
    double* own_r = D.rpart.p;
    int* own_i = D.ipart.p;
    DevStatus* sst = reinterpret_cast<DevStatus*>(base + L.st);
    HIPCHECK(hipMemcpyAsync(sst, own_st, sizeof(DevStatus), hipMemcpyDeviceToDevice, D.stream));
    const bool graphs = use_graphs_;
    struct Restore {
        Instance& I;
        Impl& D;
        DevStatus *own_st, *sst;
        double *own_m, *own_j, *own_r;
        int* own_i;
        bool graphs;
        ~Restore() {
            (void)hipMemcpyAsync(own_st, sst, sizeof(DevStatus), hipMemcpyDeviceToDevice, D.stream);
            D.st.p = own_st;
            D.mpart.p = own_m;
            D.jpart.p = own_j;
            D.rpart.p = own_r;
            D.ipart.p = own_i;
            I.use_graphs_ = graphs;
            I.bytes_.frozen = false;
        }
    } restore{*this, D, own_st, sst, own_m, own_j, own_r, own_i, graphs};
    D.st.p = sst;
    D.mpart.p = reinterpret_cast<double*>(base + L.mpart);
    if (D.m) {
        D.jpart.p = reinterpret_cast<double*>(base + L.jpart);
        D.rpart.p = reinterpret_cast<double*>(base + L.rpart);
        D.ipart.p = reinterpret_cast<int*>(base + L.ipart);
    }
    use_graphs_ = false;
    bytes_.frozen = true;
    return solve(p);
}

void Instance::get_results(double* P, double* Q, double* Pf, double* Qf, double* Pt, double* Qt, double* hx,
                           double* r) {
    HIPCHECK(hipSetDevice(device_));
    StreamScope scope_(d_->stream);
    Impl& D = *d_;
    eval_outputs(D);
    fetch_unpermute(D.P, D.lb, P, D.stream);
    fetch_unpermute(D.Q, D.lb, Q, D.stream);
    fetch_unpermute(D.Pf, D.lk, Pf, D.stream);
    fetch_unpermute(D.Qf, D.lk, Qf, D.stream);
    fetch_unpermute(D.Pt, D.lk, Pt, D.stream);
    fetch_unpermute(D.Qt, D.lk, Qt, D.stream);
    fetch_unpermute(D.hx, D.lm, hx, D.stream);
    fetch_unpermute(D.r, D.lm, r, D.stream);
}

void Instance::debug_eval(double* hx, double* P, double* Q, double* Pf, double* Qf, double* Pt, double* Qt) {
    HIPCHECK(hipSetDevice(device_));
    StreamScope scope_(d_->stream);
    Impl& D = *d_;
    launch_power(D);
    if (D.nb)
        k_flows<<<nblk(D.nb), kBlock, 0, D.stream>>>(D.nb, D.from.p, D.to.p, D.gs.p, D.bs.p, D.bc.p, D.tap.p,
                                                     D.shift.p, D.status.p, D.V.p, D.th.p, D.Pf.p, D.Qf.p, D.Pt.p,
                                                     D.Qt.p);
    launch_meas(D, 0, 1.5, 0, 0, true);
    HIPCHECK(hipGetLastError());
    get_results(P, Q, Pf, Qf, Pt, Qt, hx, nullptr);
}

void Instance::debug_ybus(double* G, double* B) {
    HIPCHECK(hipSetDevice(device_));
    StreamScope scope_(d_->stream);
    Impl& D = *d_;
    const Symbolic& S = *S_;
    const size_t nY = S.ycol.size();
    std::vector<double> g(nY), b(nY);
    HIPCHECK(hipMemcpyAsync(g.data(), D.yG.p, nY * sizeof(double), hipMemcpyDeviceToHost, D.stream));
    HIPCHECK(hipMemcpyAsync(b.data(), D.yB.p, nY * sizeof(double), hipMemcpyDeviceToHost, D.stream));
    HIPCHECK(hipStreamSynchronize(D.stream));
    for (int i = 0; i < S.n; ++i)
        for (int e = S.yrow[i]; e < S.yrow[i + 1]; ++e) {
            const int de = D.yrow_d[D.lb[i]] + (e - S.yrow[i]);
            G[e] = g[de];
            B[e] = b[de];
        }
}

namespace {
template <typename T>
void fetch_as_double(const DBuf<T>& src, size_t cnt, double* dst, hipStream_t s) {
    std::vector<T> tmp(cnt);
    if (cnt) HIPCHECK(hipMemcpyAsync(tmp.data(), src.p, cnt * sizeof(T), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    for (size_t q = 0; q < cnt; ++q) dst[q] = (double)tmp[q];
}
}  // namespace

void Instance::debug_jacobian_gain(double* hv, double* gv) {
    HIPCHECK(hipSetDevice(device_));
    StreamScope scope_(d_->stream);
    Impl& D = *d_;
    const Symbolic& S = *S_;
    launch_power(D);
    if (precision_ == 1) {
        launch_meas(D, 0, 1.5, 1, 1, false);
        launch_spgemm<double>(D, S);
        HIPCHECK(hipGetLastError());
        if (hv) {
            std::vector<double> tmp(S.hcol.size());
            fetch_as_double(D.hs, S.hcol.size(), tmp.data(), D.stream);
            for (size_t q = 0; q < S.hcol.size(); ++q) hv[q] = tmp[D.qmap[q]];
        }
        if (gv) fetch_as_double(D.f64.Gval, S.nnz_g, gv, D.stream);
    } else {
        launch_meas(D, 0, 1.5, 1, 1, false);
        launch_spgemm<float>(D, S);
        HIPCHECK(hipGetLastError());
        if (hv) {
            std::vector<double> tmp(S.hcol.size());
            fetch_as_double(D.hs, S.hcol.size(), tmp.data(), D.stream);
            for (size_t q = 0; q < S.hcol.size(); ++q) hv[q] = tmp[D.qmap[q]];
        }
        if (gv) fetch_as_double(D.f32.Gval, S.nnz_g, gv, D.stream);
    }
}

bool Instance::debug_rhs(bool by_location, double* b) {
    HIPCHECK(hipSetDevice(device_));
    StreamScope scope_(d_->stream);
    Impl& D = *d_;
    const Symbolic& S = *S_;
    if (by_location && !D.use_loc) return false;
    if (!D.f64.ready) D.f64.alloc(S, bytes_, D.pi.sizes());
    if (by_location) {
        launch_meas_loc(D, 0, 1.5);
        launch_rhs_loc<double>(D);
    } else {
        launch_power(D);
        launch_meas(D, 0, 1.5, 0, 0, false, true);
        launch_rhs<double>(D);
    }
    DBuf<double> out;
    Ledger tmp;
    out.alloc(std::max(1, D.ns), tmp);
    if (D.ns) k_permute_out<double><<<nblk(D.ns), kBlock, 0, D.stream>>>(D.ns, D.perm.p, D.f64.x.p, out.p);
    HIPCHECK(hipGetLastError());
    if (D.ns) HIPCHECK(hipMemcpyAsync(b, out.p, D.ns * sizeof(double), hipMemcpyDeviceToHost, D.stream));
    HIPCHECK(hipStreamSynchronize(D.stream));
    out.release(tmp);
    return true;
}

int Instance::debug_factor(double* fronts) {
    ++d_->plan_epoch;  // launch plans change: captured graphs retire
    HIPCHECK(hipSetDevice(device_));
    StreamScope scope_(d_->stream);
    Impl& D = *d_;
    const Symbolic& S = *S_;
    for (int attempt = 0; attempt < 2; ++attempt) {
        launch_power(D);
        if (precision_ == 1) {
            launch_meas(D, 0, 1.5, 0, 1, false);
            launch_spgemm<double>(D, S);
            launch_factor<double>(D, 1e-13);
        } else {
            launch_meas(D, 0, 1.5, 0, 1, false);
            launch_spgemm<float>(D, S);
            launch_factor<float>(D, 1e-5f);
        }
        HIPCHECK(hipGetLastError());
        read_status(D);
        if (!D.hst->flow_err) break;
        D.use_flow = false;
        D.lds_flags = false;
    }
    if (fronts) {
        if (precision_ == 1)
            fetch_as_double(D.f64.F, S.ftotal, fronts, D.stream);
        else
            fetch_as_double(D.f32.F, S.ftotal, fronts, D.stream);
    }
    factor_valid_ = D.hst->piv == 0;
    return D.hst->piv;
}

void Instance::debug_solve(const double* b, double* dx) {
    ++d_->plan_epoch;  // launch plans change: captured graphs retire
    HIPCHECK(hipSetDevice(device_));
    StreamScope scope_(d_->stream);
    Impl& D = *d_;
    DBuf<double> bb, xx;
    Ledger tmp;
    bb.alloc(D.ns, tmp);
    xx.alloc(D.ns, tmp);
    bb.upload(b, D.ns, D.stream);
    for (int attempt = 0; attempt < 2; ++attempt) {
        HIPCHECK(hipMemsetAsync(&D.st.p->flow_err, 0, sizeof(int), D.stream));
        if (precision_ == 1) {
            k_permute_in<double><<<nblk(D.ns), kBlock, 0, D.stream>>>(D.ns, D.perm.p, bb.p, D.f64.x.p);
            launch_trisolve<double>(D);
            k_permute_out<double><<<nblk(D.ns), kBlock, 0, D.stream>>>(D.ns, D.perm.p, D.f64.x.p, xx.p);
        } else {
            k_permute_in<float><<<nblk(D.ns), kBlock, 0, D.stream>>>(D.ns, D.perm.p, bb.p, D.f32.x.p);
            launch_trisolve<float>(D);
            k_permute_out<float><<<nblk(D.ns), kBlock, 0, D.stream>>>(D.ns, D.perm.p, D.f32.x.p, xx.p);
        }
        HIPCHECK(hipGetLastError());
        read_status(D);
        if (!D.hst->flow_err) break;
        D.use_flow = false;
        D.lds_flags = false;
    }
    HIPCHECK(hipMemcpyAsync(dx, xx.p, D.ns * sizeof(double), hipMemcpyDeviceToHost, D.stream));
    HIPCHECK(hipStreamSynchronize(D.stream));
    bb.release(tmp);
    xx.release(tmp);
}

void Instance::stage_telemetry(int slot, const int32_t* idx, const double* val, int count) {
    if (slot < 0 || count < 0) throw std::invalid_argument("bad staging slot or count");
    for (int q = 0; q < count; ++q)
        if (idx[q] < 0 || idx[q] >= d_->m) throw std::invalid_argument("staged measurement index out of range");
    HIPCHECK(hipSetDevice(device_));
    StreamScope scope_(d_->stream);
    Impl& D = *d_;
    if ((int)D.staged.size() <= slot) D.staged.resize(slot + 1);
    Impl::Staged& sb = D.staged[slot];
    if (sb.idx.n < (size_t)count) {
        sb.idx.alloc(count, bytes_);
        sb.val.alloc(count, bytes_);
    }
    std::vector<int32_t> di(count);
    for (int q = 0; q < count; ++q) di[q] = D.lm[idx[q]];
    sb.idx.upload(di.data(), count, D.stream);
    sb.val.upload(val, count, D.stream);
    sb.count = count;
    HIPCHECK(hipStreamSynchronize(D.stream));
}

void Instance::apply_staged(int slot) {
    Impl& D = *d_;
    if (slot < 0 || slot >= (int)D.staged.size()) throw std::invalid_argument("unknown staging slot");
    HIPCHECK(hipSetDevice(device_));
    StreamScope scope_(d_->stream);
    const Impl::Staged& sb = D.staged[slot];
    if (sb.count) keep_z(D);
    if (sb.count)
        k_scatter_z<<<nblk(sb.count), kBlock, 0, D.stream>>>(sb.count, sb.idx.p, sb.val.p, D.z.p);
    HIPCHECK(hipGetLastError());
}

void Instance::set_level_schedule(bool on) {
    ++d_->plan_epoch;  // launch plans change: captured graphs retire
    d_->force_levels = on;
    d_->use_flow = !on && d_->flow_ok;
}

bool Instance::dataflow_sweeps() const { return d_->use_flow; }

void Instance::set_factor_plan(int cut, int wcut) {
    ++d_->plan_epoch;
    d_->force_cut = cut;
    d_->force_wcut = wcut;
}

void Instance::set_task_sweeps(bool on) {
    ++d_->plan_epoch;
    d_->use_tasks = on;
}

void Instance::set_lds_solve(bool on) {
    ++d_->plan_epoch;
    d_->use_lds = on;
}

void Instance::set_dense_subtrees(bool on) {
    ++d_->plan_epoch;
    d_->pi_dense = on;
}

void Instance::set_fork_subtree_matrices(int mode) {
    ++d_->plan_epoch;  // captured refactoring iterations hold the fork or not
    d_->fork_subm = mode;
}

void Instance::set_middle_subtrees(bool on) {
    ++d_->plan_epoch;
    d_->pi_mid = on;
}

void Instance::set_subtree_packing(int mode) {
    if (mode < -1 || mode > 1) throw std::invalid_argument("subtree packing: -1 (by precision), 0 or 1");
    ++d_->plan_epoch;
    d_->sub_pack = mode;
}

void Instance::set_subtree_workgroups(int fwd, int bwd) {
    auto ok = [](int w) { return w == 0 || w == 1 || w == 4 || w == 8 || w == 16; };
    if (!ok(fwd) || !ok(bwd)) throw std::invalid_argument("waves per middle subtree: 0, 1, 4, 8 or 16");
    ++d_->plan_epoch;
    d_->sub_wg_fwd = fwd;
    d_->sub_wg_bwd = bwd;
}

void Instance::solve_plan_sizes(int64_t* out, int count) const {
    const Impl& D = *d_;
    const Impl::PiDev& P = D.pi;
    const bool on = D.use_pi && P.ok;
    const bool d = on && P.dense;
    const int64_t v[14] = {d ? P.nsub : 0,    d ? P.msub : 0, d ? P.usum : 0, on ? P.ntop : 0,  on ? P.ntc : 0,
                           on ? P.nblock : 0, on ? P.nq : 0,  on ? P.nseg : 0, on ? P.npath : 0, D.use_loc ? 1 : 0,
                           d ? P.nsub0 : 0,   d ? P.msub0 : 0, d ? P.usum0 : 0, d ? P.ncol0 : 0};
    for (int i = 0; i < count && i < 14; ++i) out[i] = v[i];
}

void Instance::set_rhs_by_location(bool on) {
    ++d_->plan_epoch;
    d_->rhs_loc = on;
}

void Instance::set_mixed_fp64_factor(bool on) {
    ++d_->plan_epoch;
    if (d_->mixed64 != on) factor_valid_ = false;
    d_->mixed64 = on;
    d_->mixed64_off = false;
}

void Instance::set_meas_by_type(bool on) {
    ++d_->plan_epoch;
    d_->meas_by_type = on;
}

void Instance::set_xcd_remap(bool on) {
    ++d_->plan_epoch;  // launch plans change: captured graphs retire
    HIPCHECK(hipSetDevice(device_));
    StreamScope scope_(d_->stream);
    const int v = on ? 1 : 0;
    HIPCHECK(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_xcd_remap), &v, sizeof(int), 0, hipMemcpyHostToDevice, d_->stream));
    HIPCHECK(hipStreamSynchronize(d_->stream));
}

void Instance::set_partitioned_inverse(bool on, int cut, int waves) {
    ++d_->plan_epoch;  // launch plans change: captured graphs retire
    if (d_->use_pi != on) factor_valid_ = false;  // the inverse blocks are made with the factor
    d_->use_pi = on;
    if (cut >= 0) d_->pi_cut = cut;
    if (waves >= 1 && waves <= 16) d_->pi_waves = waves;
}

void Instance::set_lds_flags(bool on) {
    ++d_->plan_epoch;
    d_->lds_flags = on;
}

bool Instance::lds_flags() const { return d_->lds_flags; }

void Instance::set_lds_waves(int bottom, int tail) {
    ++d_->plan_epoch;  // launch plans change: captured graphs retire
    if (bottom >= 1 && bottom <= 16) d_->lds_waves_bottom = bottom;
    if (tail >= 1 && tail <= 16) d_->lds_waves_tail = tail;
}

std::string Instance::sweep_plan() const {
    const Impl& D = *d_;
    char buf[512];
    if (D.use_flow) return "dependency-driven";
    std::string out;
    if (!(D.use_tasks && D.tasks_ok)) {
        out = "factor levels(" + std::to_string(D.n_levels) + ")";
    } else {
        snprintf(buf, sizeof(buf), "factor tasks(cut=%d bottom=%d mid=%d tail=%d est=%.0fus vs levels=%.0fus)",
                 D.plan_cut, D.n_bottom, (int)D.mid_off.size(), D.has_tail ? 1 : 0, D.plan_cost_us, D.level_cost_us);
        out = buf;
    }
    const Impl::LdsDev& L = D.lds[precision_ == 1 || fp64_factor_active_ ? 1 : 0];
    if (D.use_pi && D.pi.ok && !D.use_flow) {
        const Impl::LdsDev& PB = D.pi.bottom[precision_ == 1 || fp64_factor_active_ ? 1 : 0];
        int forked = 0;  // subtree matrices a factorization forms on the second stream
        if (D.fork_subm > 0 && sub_m_forks(D))
            for (int q = 0; q < kSubBins; ++q) forked += D.pi.bin_early[q];
        if (D.pi.dense)
            snprintf(buf, sizeof(buf),
                     "; solve partitioned-inverse(cut=%d dense-subtrees=%d+%d/%lld entries top supernodes=%d "
                     "columns=%d blocks=%lld row-entries=%lld segments=%d; bottom rows<=16/32/64/128 %d/%d/%d/%d "
                     "cols<=8/16/32/64 %d/%d/%d/%d; middle rows %d/%d/%d/%d cols %d/%d/%d/%d; forked matrices=%d)",
                     D.pi.cut, D.pi.nsub0, D.pi.nsub - D.pi.nsub0, (long long)D.pi.msub, D.pi.ntop, D.pi.ntc,
                     (long long)D.pi.nblock, (long long)D.pi.nq, D.pi.nseg, D.pi.rhist[0][0], D.pi.rhist[0][1],
                     D.pi.rhist[0][2], D.pi.rhist[0][3], D.pi.nhist[0][0], D.pi.nhist[0][1], D.pi.nhist[0][2],
                     D.pi.nhist[0][3], D.pi.rhist[1][0], D.pi.rhist[1][1], D.pi.rhist[1][2], D.pi.rhist[1][3],
                     D.pi.nhist[1][0], D.pi.nhist[1][1], D.pi.nhist[1][2], D.pi.nhist[1][3], forked);
        else
            snprintf(buf, sizeof(buf),
                     "; solve partitioned-inverse(cut=%d bottom=%d/%zuB%s top supernodes=%d columns=%d blocks=%lld "
                     "row-entries=%lld segments=%d)",
                     D.pi.cut, PB.n_bottom, PB.bottom_bytes, D.lds_flags ? " flags" : "", D.pi.ntop, D.pi.ntc,
                     (long long)D.pi.nblock, (long long)D.pi.nq, D.pi.nseg);
        out += buf;
    } else if (D.use_lds && L.ok) {
        snprintf(buf, sizeof(buf), "; solve lds-tasks(cut=%d bottom=%d/%zuB mid=%d tail-chunks=%zu est=%.0fus)", L.cut,
                 L.n_bottom, L.bottom_bytes, (int)L.mid_off.size(), L.tail_tasks.size(), L.cost_us);
        out += buf;
    } else {
        out += "; solve as factor";
    }
    return out;
}

namespace {
// Average time of one forward + backward solve with the current plan (the
// right-hand side is restored before every repetition).
template <typename T>
double time_solves(Instance::Impl& D, Ledger& bytes, int reps) {
    PrecBufs<T>& B = bufs<T>(D);
    B.alloc_pi(D.pi.sizes(), bytes);
    launch_solve_prep<T>(D);
    DBuf<T> save;
    Ledger tmp;
    save.alloc(B.x.n, tmp);
    HIPCHECK(hipMemcpyAsync(save.p, B.x.p, B.x.n * sizeof(T), hipMemcpyDeviceToDevice, D.stream));
    launch_trisolve<T>(D);  // warm-up
    HIPCHECK(hipEventRecord(D.ev0, D.stream));
    for (int r = 0; r < reps; ++r) {
        HIPCHECK(hipMemcpyAsync(B.x.p, save.p, B.x.n * sizeof(T), hipMemcpyDeviceToDevice, D.stream));
        launch_trisolve<T>(D);
    }
    HIPCHECK(hipEventRecord(D.ev1, D.stream));
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipEventSynchronize(D.ev1));
    float ms = 0;
    HIPCHECK(hipEventElapsedTime(&ms, D.ev0, D.ev1));
    HIPCHECK(hipMemcpyAsync(B.x.p, save.p, B.x.n * sizeof(T), hipMemcpyDeviceToDevice, D.stream));
    HIPCHECK(hipStreamSynchronize(D.stream));
    save.release(tmp);
    return ms / std::max(1, reps);
}
// Average time of the factorization sweeps alone (G unchanged between repetitions).
template <typename T>
double time_factor(Instance::Impl& D, int reps) {
    launch_factor_sweeps<T>(D, sizeof(T) == sizeof(double) ? T(1e-13) : T(1e-5));  // warm-up
    HIPCHECK(hipEventRecord(D.ev0, D.stream));
    for (int r = 0; r < reps; ++r) launch_factor_sweeps<T>(D, sizeof(T) == sizeof(double) ? T(1e-13) : T(1e-5));
    HIPCHECK(hipEventRecord(D.ev1, D.stream));
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipEventSynchronize(D.ev1));
    float ms = 0;
    HIPCHECK(hipEventElapsedTime(&ms, D.ev0, D.ev1));
    return ms / std::max(1, reps);
}
}  // namespace

std::string Instance::autotune_solve(int reps) {
    ++d_->plan_epoch;  // launch plans change: captured graphs retire
    HIPCHECK(hipSetDevice(device_));
    StreamScope scope_(d_->stream);
    Impl& D = *d_;
    const Symbolic& S = *S_;
    if (!factor_valid_) throw std::invalid_argument("autotune needs a factored gain matrix: solve first");
    if (!D.use_pi || D.use_flow) return "partitioned inverse not in use";
    const bool f64 = precision_ == 1 || fp64_factor_active_;
    // mixed form on the FP64 factor: FP32 solves, FP64 factorization
    const bool split = !f64 && D.mixed64;
    // dense bottom subtrees have no in-task schedule: only the cut matters
    const bool dense = D.pi_dense;
    const std::vector<int> cuts = dense ? std::vector<int>{8, 16, 32, 64, 128} : std::vector<int>{16, 32, 64, 128};
    const std::vector<int> waves = dense ? std::vector<int>{D.pi_waves} : std::vector<int>{2, 4, 8};
    const int nflags = dense ? 1 : 2;
    int bc = D.pi_cut, bw = D.pi_waves;
    bool bf = D.lds_flags;
    double best = 1e300;
    std::string log;
    char buf[96];
    for (int cut : cuts)
        for (int w : waves) {
            D.pi_cut = cut;
            D.pi_waves = w;
            build_pi(D, S, bytes_, D.stream);
            if (!D.pi.ok) continue;
            for (int fl = 0; fl < nflags; ++fl) {
                D.lds_flags = dense ? D.lds_flags : fl != 0;
                const double ms = f64 ? time_solves<double>(D, bytes_, reps) : time_solves<float>(D, bytes_, reps);
                snprintf(buf, sizeof(buf), "%s%d/%d%s:%.3f", log.empty() ? "" : " ", cut, w, fl ? "f" : "", ms);
                log += buf;
                if (ms < best) {
                    best = ms;
                    bc = cut;
                    bw = w;
                    bf = fl != 0;
                }
            }
        }
    D.pi_cut = bc;
    D.pi_waves = bw;
    D.lds_flags = bf;
    build_pi(D, S, bytes_, D.stream);
    if (D.f32.ready) D.f32.alloc_pi(D.pi.sizes(), bytes_);
    if (D.f64.ready) D.f64.alloc_pi(D.pi.sizes(), bytes_);
    // factorization plan: time the task sweeps of each candidate on the current G
    std::string flog;
    if (D.tasks_ok && D.use_tasks) {
        int fc = D.plan_cut, fw = D.plan_wcut;
        double fbest = 1e300;
        for (int cut : {64, 128, 256, 512, 1024})
            for (int wc : {4, 8, 16, 32, 64}) {
                build_factor_plan(D, S, cut, wc, bytes_, D.stream);
                const double ms = f64 || split ? time_factor<double>(D, reps) : time_factor<float>(D, reps);
                snprintf(buf, sizeof(buf), "%s%d/%d:%.3f", flog.empty() ? "" : " ", cut, wc, ms);
                flog += buf;
                if (ms < fbest) {
                    fbest = ms;
                    fc = cut;
                    fw = wc;
                }
            }
        build_factor_plan(D, S, fc, fw, bytes_, D.stream);
        snprintf(buf, sizeof(buf), "; factor cut=%d wcut=%d (%.3f ms) cut/wcut:ms ", fc, fw, fbest);
        flog = buf + flog;
    }
    // the factor, solve matrices and top inverse of the chosen plans
    if (f64 || split) {
        launch_factor<double>(D, 1e-13);
        if (split) round_solve_data(D, bytes_);
    } else {
        launch_factor<float>(D, 1e-5f);
    }
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(D.stream));
    snprintf(buf, sizeof(buf), "solve cut=%d waves=%d flags=%d (%.3f ms per solve) ", bc, bw, bf ? 1 : 0, best);
    return buf + std::string("cut/waves[f]:ms ") + log + flog;
}

int Instance::partitioned_inverse_cut() const { return d_->pi_cut > 0 ? d_->pi_cut : kPiCutDefault; }
int Instance::partitioned_inverse_waves() const { return d_->pi_waves; }

int Instance::staged_count(int slot) const {
    if (slot < 0 || slot >= (int)d_->staged.size()) return -1;
    return d_->staged[slot].count;
}

int device_count() {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) return 0;
    return c;
}

int runtime_version() {
    int v = 0;
    (void)hipRuntimeGetVersion(&v);
    return v;
}

const char* device_name(int dev) {
    static thread_local char buf[256];
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return "unavailable";
    snprintf(buf, sizeof(buf), "%s (%s), %d CUs, %.1f GiB HBM", prop.name, prop.gcnArchName,
             prop.multiProcessorCount, prop.totalGlobalMem / (1024.0 * 1024.0 * 1024.0));
    return buf;
}

}  // namespace sle
