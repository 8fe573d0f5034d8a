// Learned proposal width: a 3 -> 32 -> 32 -> 1 tanh MLP trained on the GPU by these kernels.
//
// The network maps features of a site's local environment that the site's own move cannot change,
//   f0 = |v| beta       (v: exchange + DMI + field local field of the neighbours)
//   f1 = K beta          (K = |Du| + |Ku_i| + |K1_i|: anisotropy strength at the site)
//   f2 = |v . e| / |v|   (alignment of the local field with the anisotropy axis e)
// to the half-angle w = WMIN + (pi - WMIN) sigmoid(o) of a cone proposal around the current spin. Because w is
// a function of those features only, the reverse move sees the same w and the proposal stays symmetric.
//
// Objective: the expected squared jump distance per update, J = E_t[min(1, exp(-beta dE(s -> t))) |t - s|^2], for
// trial spins t = T(s, w, u1, u2) of the cone proposal, estimated with K draws (u1, u2) per sample, so the
// gradient flows through w by reparameterisation. Training maximises the mean of log(1e-4 + J) over samples
// taken from the sampler's own warm-up sweeps (sck_amd.proposal).
//
// Kernels: sck_mlp_fwdbwd_k (per sample and draw: forward, objective, backward to the layer deltas),
// sck_mlp_reduce_k (per-parameter partial sums over sample chunks, fixed order), sck_mlp_adam_k (sum the partials,
// Adam step), sck_mlp_table_k (the frozen network on the grid of quantised features the sweep kernel reads).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sck.h"

#define H 32
#define NIN 3
#define NPARAM (H * NIN + H + H * H + H + H + 1)
// parameter layout: W1[H][NIN], b1[H], W2[H][H], b2[H], W3[H], b3
#define OFF_W1 0
#define OFF_B1 (OFF_W1 + H * NIN)
#define OFF_W2 (OFF_B1 + H)
#define OFF_B2 (OFF_W2 + H * H)
#define OFF_W3 (OFF_B2 + H)
#define OFF_B3 (OFF_W3 + H)

namespace {

__device__ __forceinline__ void philox10(uint32_t &c0, uint32_t &c1, uint32_t &c2, uint32_t &c3, uint32_t k0,
                                         uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
        const uint64_t p2 = (uint64_t)0xD2511F53u * (uint64_t)c0;
        const uint32_t n0 = (uint32_t)(p0 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = (uint32_t)(p2 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p0;
        c3 = (uint32_t)p2;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

__device__ __forceinline__ float u01(uint32_t x) {
    int32_t v = (int32_t)x;
    v = v < 0 ? ~v : v;
    return (float)v * 4.6566127342e-10f;
}

__device__ __forceinline__ float sigmoidf_(float x) { return 1.0f / (1.0f + expf(-x)); }

// forward pass of the network; keeps the hidden activations for the backward pass
__device__ float mlp_forward(const float *__restrict__ P, const float x[NIN], float h1[H], float h2[H]) {
    for (int j = 0; j < H; ++j) {
        float a = P[OFF_B1 + j];
        for (int i = 0; i < NIN; ++i) a += P[OFF_W1 + j * NIN + i] * x[i];
        h1[j] = tanhf(a);
    }
    for (int j = 0; j < H; ++j) {
        float a = P[OFF_B2 + j];
        for (int i = 0; i < H; ++i) a += P[OFF_W2 + j * H + i] * h1[i];
        h2[j] = tanhf(a);
    }
    float o = P[OFF_B3];
    for (int i = 0; i < H; ++i) o += P[OFF_W3 + i] * h2[i];
    return o;
}

__device__ __forceinline__ float dotn(const float *a, const float *b, int n) {
    float d = 0.0f;
    for (int k = 0; k < n; ++k) d += a[k] * b[k];
    return d;
}

// energy change of s -> t (vector models) and its gradient with respect to t
__device__ float delta_e(const sck_mlp_sample &S, const sck_mlp_config &cfg, const float *t, float *g) {
    const int n = cfg.n;
    float dE = 0.0f;
    for (int k = 0; k < n; ++k) {
        dE -= (t[k] - S.s[k]) * S.v[k];
        g[k] = -S.v[k];
    }
    if (n == 3 && (S.k1 != 0.0f || cfg.k2 != 0.0f)) {
        // cubic: a(p) = k1 ((xy)^2 + (xz)^2 + (yz)^2) + k2 (xyz)^2
        const float *p = t;
        const float *q = S.s;
        const float at = S.k1 * (p[0] * p[0] * p[1] * p[1] + p[0] * p[0] * p[2] * p[2] + p[1] * p[1] * p[2] * p[2]) +
                         cfg.k2 * p[0] * p[0] * p[1] * p[1] * p[2] * p[2];
        const float as = S.k1 * (q[0] * q[0] * q[1] * q[1] + q[0] * q[0] * q[2] * q[2] + q[1] * q[1] * q[2] * q[2]) +
                         cfg.k2 * q[0] * q[0] * q[1] * q[1] * q[2] * q[2];
        dE += at - as;
        g[0] += 2.0f * S.k1 * p[0] * (p[1] * p[1] + p[2] * p[2]) + 2.0f * cfg.k2 * p[0] * p[1] * p[1] * p[2] * p[2];
        g[1] += 2.0f * S.k1 * p[1] * (p[0] * p[0] + p[2] * p[2]) + 2.0f * cfg.k2 * p[1] * p[0] * p[0] * p[2] * p[2];
        g[2] += 2.0f * S.k1 * p[2] * (p[0] * p[0] + p[1] * p[1]) + 2.0f * cfg.k2 * p[2] * p[0] * p[0] * p[1] * p[1];
    }
    if (cfg.du != 0.0f) {
        const float te = dotn(t, cfg.axis, n), se = dotn(S.s, cfg.axis, n);
        dE += cfg.du * (se * se - te * te);
        for (int k = 0; k < n; ++k) g[k] -= 2.0f * cfg.du * te * cfg.axis[k];
    }
    if (S.ku != 0.0f) {
        const float tn = dotn(t, S.axis, n), sn = dotn(S.s, S.axis, n);
        dE += S.ku * (sn * sn - tn * tn);
        for (int k = 0; k < n; ++k) g[k] -= 2.0f * S.ku * tn * S.axis[k];
    }
    return dE;
}

// trial spin of the cone proposal of half-angle w and dt/dw
__device__ void cone(const sck_mlp_config &cfg, const float *s, float w, float u1, float u2, float *t, float *dt) {
    if (cfg.n == 2) {
        const float r = 2.0f * u1 - 1.0f;
        const float ph = atan2f(s[1], s[0]) + w * r;
        float sp, cp;
        sincosf(ph, &sp, &cp);
        t[0] = cp;
        t[1] = sp;
        dt[0] = -sp * r;
        dt[1] = cp * r;
        return;
    }
    float sw, cw;
    sincosf(w, &sw, &cw);
    const float c = 1.0f - u1 * (1.0f - cw);
    const float dc = -u1 * sw;
    const float st = sqrtf(fmaxf(1.0f - c * c, 1e-30f));
    const float dst = -c * dc / st;
    const float phi = 6.283185307179586f * u2;
    float sp, cp;
    sincosf(phi, &sp, &cp);
    const bool nz = fabsf(s[2]) > 0.9f;
    const float ax0 = nz ? 1.0f : 0.0f, ax1 = 0.0f, ax2 = nz ? 0.0f : 1.0f;
    float e1[3] = {ax1 * s[2] - ax2 * s[1], ax2 * s[0] - ax0 * s[2], ax0 * s[1] - ax1 * s[0]};
    const float ne = sqrtf(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
    for (int k = 0; k < 3; ++k) e1[k] /= ne;
    const float e2[3] = {s[1] * e1[2] - s[2] * e1[1], s[2] * e1[0] - s[0] * e1[2], s[0] * e1[1] - s[1] * e1[0]};
    for (int k = 0; k < 3; ++k) {
        const float perp = cp * e1[k] + sp * e2[k];
        t[k] = c * s[k] + st * perp;
        dt[k] = dc * s[k] + dst * perp;
    }
}

}  // namespace

// Per sample: features -> network -> w, then ndraw cone draws give the sample's expected squared jump
// J_i = mean_k min(1, exp(-beta dE)) |t_k - s|^2 and dJ_i/dw. The loss is L = -mean_i log(EPS_J + J_i): a
// sample's weight does not depend on the scale of its jumps, so cold and hot temperatures count alike. The
// backward pass writes one record per sample:
//   rec[0..2] = x, rec[3..3+H) = h1, rec[3+H..3+2H) = h2, rec[3+2H..3+3H) = delta1, rec[3+3H..3+4H) = delta2,
//   rec[3+4H] = delta3 (= dL/do), rec[4+4H] = J_i.
#define EPS_J 1e-4f
extern "C" __global__ void __launch_bounds__(128) sck_mlp_fwdbwd_k(const float *__restrict__ P,
                                                                 const sck_mlp_sample *__restrict__ samples,
                                                                 int32_t nsample, sck_mlp_config cfg,
                                                                 float *__restrict__ rec) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nsample) return;
    const sck_mlp_sample S = samples[i];
    float x[NIN];
    const int n = cfg.n;
    const float hv = sqrtf(dotn(S.v, S.v, n));
    x[0] = hv * S.beta * cfg.xs[0];
    x[1] = (fabsf(cfg.du) + fabsf(S.ku) + fabsf(S.k1)) * S.beta * cfg.xs[1];
    x[2] = (hv > 0.0f ? fabsf(dotn(S.v, cfg.axis, n)) / hv : 0.0f) * cfg.xs[2];
    float h1[H], h2[H];
    const float o = mlp_forward(P, x, h1, h2);
    const float sg = sigmoidf_(o);
    const float w = cfg.wmin + (cfg.wmax - cfg.wmin) * sg;
    const float dwdo = (cfg.wmax - cfg.wmin) * sg * (1.0f - sg);
    float Ji = 0.0f, dJdw = 0.0f;
    for (int k = 0; k < cfg.ndraw; ++k) {
        uint32_t c0 = (uint32_t)(i * cfg.ndraw + k), c1 = 0u, c2 = 0u, c3 = 0u;
        philox10(c0, c1, c2, c3, (uint32_t)cfg.seed, 0u);
        const float u1 = u01(c0), u2 = u01(c1);
        float t[3] = {0, 0, 0}, dt[3] = {0, 0, 0}, g[3] = {0, 0, 0};
        cone(cfg, S.s, w, u1, u2, t, dt);
        const float dE = delta_e(S, cfg, t, g);
        const float xb = -S.beta * dE;
        const float a = xb >= 0.0f ? 1.0f : expf(xb);
        const float dadE = xb >= 0.0f ? 0.0f : -S.beta * a;
        float d = 0.0f;
        for (int q = 0; q < n; ++q) d += (t[q] - S.s[q]) * (t[q] - S.s[q]);
        Ji += a * d;
        for (int q = 0; q < n; ++q) dJdw += (d * dadE * g[q] + a * 2.0f * (t[q] - S.s[q])) * dt[q];
    }
    Ji /= (float)cfg.ndraw;
    dJdw /= (float)cfg.ndraw;
    const float d3 = -dJdw / (EPS_J + Ji) * dwdo / (float)nsample;
    float *r = rec + (int64_t)i * (5 + 4 * H);
    for (int q = 0; q < NIN; ++q) r[q] = x[q];
    for (int j = 0; j < H; ++j) {
        r[3 + j] = h1[j];
        r[3 + H + j] = h2[j];
    }
    float d2[H];
    for (int j = 0; j < H; ++j) {
        d2[j] = d3 * P[OFF_W3 + j] * (1.0f - h2[j] * h2[j]);
        r[3 + 3 * H + j] = d2[j];
    }
    for (int j = 0; j < H; ++j) {
        float sacc = 0.0f;
        for (int m = 0; m < H; ++m) sacc += P[OFF_W2 + m * H + j] * d2[m];
        r[3 + 2 * H + j] = sacc * (1.0f - h1[j] * h1[j]);
    }
    r[3 + 4 * H] = d3;
    r[4 + 4 * H] = Ji;
}

// Partial gradient sums: blockIdx.x = chunk of records, one thread per parameter (NPARAM + 1 threads; the
// last one sums the objective), fixed summation order.
extern "C" __global__ void __launch_bounds__(1024) sck_mlp_reduce_k(const float *__restrict__ rec, int32_t nrec,
                                                                  int32_t chunk, float *__restrict__ partial) {
    const int p = threadIdx.x + blockIdx.y * blockDim.x;
    if (p > NPARAM) return;
    const int lo = blockIdx.x * chunk, hi = min(lo + chunk, nrec);
    const int RS = 5 + 4 * H;
    float acc = 0.0f;
    for (int i = lo; i < hi; ++i) {
        const float *r = rec + (int64_t)i * RS;
        float gv;
        if (p < OFF_B1) {                       // W1[j][q] = delta1[j] x[q]
            const int j = p / NIN, q = p - j * NIN;
            gv = r[3 + 2 * H + j] * r[q];
        } else if (p < OFF_W2) {                // b1[j]
            gv = r[3 + 2 * H + (p - OFF_B1)];
        } else if (p < OFF_B2) {                // W2[j][m] = delta2[j] h1[m]
            const int j = (p - OFF_W2) / H, m = (p - OFF_W2) - j * H;
            gv = r[3 + 3 * H + j] * r[3 + m];
        } else if (p < OFF_W3) {                // b2[j]
            gv = r[3 + 3 * H + (p - OFF_B2)];
        } else if (p < OFF_B3) {                // W3[m] = delta3 h2[m]
            gv = r[3 + 4 * H] * r[3 + H + (p - OFF_W3)];
        } else if (p == OFF_B3) {               // b3
            gv = r[3 + 4 * H];
        } else {                                // the objective log(EPS_J + J_i) (for monitoring)
            gv = logf(EPS_J + r[4 + 4 * H]);
        }
        acc += gv;
    }
    partial[(int64_t)blockIdx.x * (NPARAM + 1) + p] = acc;
}

// Sum the partials (fixed order) into grad[NPARAM + 1], then one Adam step on the parameters.
extern "C" __global__ void __launch_bounds__(256) sck_mlp_adam_k(float *__restrict__ P, float *__restrict__ m,
                                                               float *__restrict__ v, float *__restrict__ grad,
                                                               const float *__restrict__ partial, int32_t nchunk,
                                                               sck_mlp_adam_config ac) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > NPARAM) return;
    float g = 0.0f;
    for (int c = 0; c < nchunk; ++c) g += partial[(int64_t)c * (NPARAM + 1) + p];
    grad[p] = g;
    if (p == NPARAM || ac.lr == 0.0f) return;
    const float mm = ac.beta1 * m[p] + (1.0f - ac.beta1) * g;
    const float vv = ac.beta2 * v[p] + (1.0f - ac.beta2) * g * g;
    m[p] = mm;
    v[p] = vv;
    const float mh = mm / (1.0f - powf(ac.beta1, (float)ac.step));
    const float vh = vv / (1.0f - powf(ac.beta2, (float)ac.step));
    P[p] -= ac.lr * mh / (sqrtf(vh) + ac.eps);
}

// The frozen network on the grid of quantised features: table[i0][i1][i2] at the bin centres
// f = (i + 0.5) / s (s: the grid's bins per unit of feature).
extern "C" __global__ void __launch_bounds__(128) sck_mlp_table_k(const float *__restrict__ P, sck_mlp_config cfg,
                                                                sck_width_grid grid, float *__restrict__ table) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    const int total = grid.n[0] * grid.n[1] * grid.n[2];
    if (id >= total) return;
    const int i2 = id % grid.n[2], i1 = (id / grid.n[2]) % grid.n[1], i0 = id / (grid.n[1] * grid.n[2]);
    float x[NIN];
    x[0] = ((float)i0 + 0.5f) / grid.s[0] * cfg.xs[0];
    x[1] = ((float)i1 + 0.5f) / grid.s[1] * cfg.xs[1];
    x[2] = ((float)i2 + 0.5f) / grid.s[2] * cfg.xs[2];
    float h1[H], h2[H];
    const float o = mlp_forward(P, x, h1, h2);
    table[id] = cfg.wmin + (cfg.wmax - cfg.wmin) * sigmoidf_(o);
}
