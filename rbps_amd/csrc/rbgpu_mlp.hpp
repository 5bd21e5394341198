// Learned warm start of GJK/EPA: a small MLP (32 -> 128 -> 128 -> 128 -> 16, ReLU) on
// BF16 MFMA with FP32 accumulation, with its own inference, training (forward and
// backward) and Adam kernels.  It maps the relative pose of a GJK pair plus a
// descriptor of both shapes to a GJK initial search direction and an EPA seed normal,
// both in the frame of shape A.  The reference has no counterpart: its GJK seed is
// tf_b.pos - tf_a.pos (ContactManifoldGenerator.cpp:423-425) and EPA takes no seed
// (EPA.cpp:152-163).  Plain GJK/EPA still run to their termination tolerances and are
// the fallback, so the prediction changes iteration counts, not the tolerance.
//
// Features (32): relative position of B in A's frame (3), relative rotation matrix
// conj(qA) qB (9, column-major), per shape the centre and half extents of its local
// bounds (6 + 6), shape kind one-hot sphere/box/capsule/hull per shape (4 + 4).
// Outputs: 0-2 GJK direction, 3-5 EPA seed normal (normalised by the consumer); 6-15 pad.
// Loss per sample: (1 - cos(out[0:3], t_gjk)) + m_epa (1 - cos(out[3:6], t_epa)), t_gjk =
// the EPA normal of an overlapping pair or the reverse of the GJK exit direction of a
// separated one (GJK's first search direction is -support(seed): a seed pointing away
// from the origin makes it cross the Minkowski difference toward the origin, which ends
// a separated pair's GJK after one step far more often than the reference's seed
// tf_b.pos - tf_a.pos does), t_epa = the EPA normal (m_epa = 1 when EPA converged).
//
// MFMA: v_mfma_f32_16x16x32_bf16 (CDNA4).  Lane l holds A[l % 16][8 (l / 16) + i] and
// B[8 (l / 16) + i][l % 16] for i < 8, and C[4 (l / 16) + j][l % 16] for j < 4.  Operands
// are gathered element by element through strides, so one routine serves the forward
// products, the input-gradient products and the weight-gradient products (contraction
// over the samples of a tile).
#pragma once
#include "rbgpu_gjk.hpp"

namespace rb {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int MLP_IN = 32, MLP_H = 128, MLP_OUT = 16;
constexpr int MLP_TILE = 64;                  // samples per workgroup step: 4 waves x 16
// flat parameter layout: W[out][in] then bias, per layer
constexpr int MLP_W1 = 0, MLP_B1 = MLP_W1 + MLP_H * MLP_IN;
constexpr int MLP_W2 = MLP_B1 + MLP_H, MLP_B2 = MLP_W2 + MLP_H * MLP_H;
constexpr int MLP_W3 = MLP_B2 + MLP_H, MLP_B3 = MLP_W3 + MLP_H * MLP_H;
constexpr int MLP_W4 = MLP_B3 + MLP_H, MLP_B4 = MLP_W4 + MLP_OUT * MLP_H;
constexpr int MLP_NPARAM = MLP_B4 + MLP_OUT;  // 39,312

struct MlpParams {
    const bf16 *wb;    // BF16 copy of the parameters (weights used from here)
    const float *wf;   // FP32 master parameters (biases used from here)
};

__device__ __forceinline__ float bf2f(bf16 x) { return (float)x; }
__device__ __forceinline__ bf16 f2bf(float x) { return (bf16)x; }

// acc += A(16 x K) B(K x 16), A(m, k) = a[m sam + k sak], B(k, n) = b[k sbk + n sbn];
// k >= K reads as zero.
template <class TA, class TB>
__device__ __forceinline__ floatx4 mfma16(const TA *a, int sam, int sak, const TB *b, int sbk, int sbn, int K, floatx4 acc) {
    const int l = lane64(), r = l & 15, kq = (l >> 4) * 8;
    for (int k0 = 0; k0 < K; k0 += 32) {
        bf16x8 av, bv;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int k = k0 + kq + i;
            av[i] = k < K ? (bf16)a[r * sam + k * sak] : (bf16)0.0f;
            bv[i] = k < K ? (bf16)b[k * sbk + r * sbn] : (bf16)0.0f;
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc, 0, 0, 0);
    }
    return acc;
}

// acc += A(16 x K) W(16 x K)^T with both operands row-major and K a multiple of 32: each
// lane's eight K values are contiguous, one 16-byte load per operand (rows 16-byte aligned).
__device__ __forceinline__ floatx4 mfma16_rows(const bf16 *a, int lda, const bf16 *w, int ldw, int K, floatx4 acc) {
    const int l = lane64(), r = l & 15, kq = (l >> 4) * 8;
    for (int k0 = 0; k0 < K; k0 += 32) {
        const bf16x8 av = *reinterpret_cast<const bf16x8 *>(a + r * lda + k0 + kq);
        const bf16x8 bv = *reinterpret_cast<const bf16x8 *>(w + r * ldw + k0 + kq);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc, 0, 0, 0);
    }
    return acc;
}

// One dense layer for a wave's 16 rows: out(s, n) = act(sum_k in(s, k) W(n, k) + b(n)).
// in: [16][ld_in] (this wave's rows), W: [N][K] row-major, out: [16][ld_out].
template <bool RELU, class TO>
__device__ __forceinline__ void dense16(const bf16 *in, int ld_in, int K, const bf16 *W, const float *bias, int N, TO *out,
                                        int ld_out) {
    const int l = lane64();
    for (int nb = 0; nb < N; nb += 16) {
        floatx4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        acc = mfma16_rows(in, ld_in, W + nb * K, K, K, acc);
        const int n = nb + (l & 15);
        const float bn = bias[n];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            float z = acc[j] + bn;
            if (RELU) z = z > 0.0f ? z : 0.0f;
            out[(4 * (l >> 4) + j) * ld_out + n] = (TO)z;
        }
    }
}

// The same layer for RT row tiles of 16 (rows 16 t + ..): each weight fragment is loaded
// once and used for every tile; each output sees the same MFMA sequence as in dense16.
template <bool RELU, int RT, class TO>
__device__ __forceinline__ void dense_rows(const bf16 *in, int ld_in, int K, const bf16 *W, const float *bias, int N,
                                           TO *out, int ld_out) {
    const int l = lane64(), r = l & 15, kq = (l >> 4) * 8;
    for (int nb = 0; nb < N; nb += 16) {
        floatx4 acc[RT];
#pragma unroll
        for (int t = 0; t < RT; ++t) acc[t] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
        for (int k0 = 0; k0 < K; k0 += 32) {
            const bf16x8 bv = *reinterpret_cast<const bf16x8 *>(W + (nb + r) * K + k0 + kq);
#pragma unroll
            for (int t = 0; t < RT; ++t) {
                const bf16x8 av = *reinterpret_cast<const bf16x8 *>(in + (16 * t + r) * ld_in + k0 + kq);
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[t], 0, 0, 0);
            }
        }
        const int n = nb + r;
        const float bn = bias[n];
#pragma unroll
        for (int t = 0; t < RT; ++t)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                float z = acc[t][j] + bn;
                if (RELU) z = z > 0.0f ? z : 0.0f;
                out[(16 * t + 4 * (l >> 4) + j) * ld_out + n] = (TO)z;
            }
    }
}

// Forward of a wave's 16 samples: x [16][32] -> a1, a2, a3 [16][128] (BF16) -> out [16][16].
__device__ __forceinline__ void mlp_forward16(const MlpParams &P, const bf16 *x, bf16 *a1, bf16 *a2, bf16 *a3, float *out) {
    dense16<true>(x, MLP_IN, MLP_IN, P.wb + MLP_W1, P.wf + MLP_B1, MLP_H, a1, MLP_H);
    __syncthreads();
    dense16<true>(a1, MLP_H, MLP_H, P.wb + MLP_W2, P.wf + MLP_B2, MLP_H, a2, MLP_H);
    __syncthreads();
    dense16<true>(a2, MLP_H, MLP_H, P.wb + MLP_W3, P.wf + MLP_B3, MLP_H, a3, MLP_H);
    __syncthreads();
    dense16<false>(a3, MLP_H, MLP_H, P.wb + MLP_W4, P.wf + MLP_B4, MLP_OUT, out, MLP_OUT);
    __syncthreads();
}

// Inference over n samples (features BF16 [n][32]); out: per sample two float4, the
// normalised GJK direction and EPA seed normal (zero vectors when degenerate).  Two
// wavefronts of MLP_INFER_THREADS, each with two 16-sample tiles of the block's MLP_TILE.
constexpr int MLP_INFER_RT = 2, MLP_INFER_THREADS = 64 * MLP_TILE / (16 * MLP_INFER_RT);
__global__ void __launch_bounds__(MLP_INFER_THREADS) k_mlp_infer(const bf16 *feat, uint32_t n_max, const uint32_t *n_dev,
                                                                 MlpParams P, float4 *out) {
    __shared__ __attribute__((aligned(16))) bf16 x[MLP_TILE * MLP_IN], h1[MLP_TILE * MLP_H], h2[MLP_TILE * MLP_H];
    __shared__ float o[MLP_TILE * MLP_OUT];
    const int wave = threadIdx.x >> 6, l = lane64();
    const uint32_t n = n_dev ? min(*n_dev, n_max) : n_max;   // the count may live on the device
    for (uint32_t t0 = blockIdx.x * MLP_TILE; t0 < n; t0 += gridDim.x * MLP_TILE) {
        for (int i = threadIdx.x; i < MLP_TILE * MLP_IN; i += blockDim.x) {
            const uint32_t s = t0 + i / MLP_IN;
            x[i] = s < n ? feat[(size_t)t0 * MLP_IN + i] : (bf16)0.0f;
        }
        __syncthreads();
        constexpr int RT = MLP_INFER_RT;
        const int r0 = wave * 16 * RT;
        bf16 *xa = x + r0 * MLP_IN, *a1 = h1 + r0 * MLP_H, *a2 = h2 + r0 * MLP_H;
        dense_rows<true, RT>(xa, MLP_IN, MLP_IN, P.wb + MLP_W1, P.wf + MLP_B1, MLP_H, a1, MLP_H);
        __syncthreads();
        dense_rows<true, RT>(a1, MLP_H, MLP_H, P.wb + MLP_W2, P.wf + MLP_B2, MLP_H, a2, MLP_H);
        __syncthreads();
        dense_rows<true, RT>(a2, MLP_H, MLP_H, P.wb + MLP_W3, P.wf + MLP_B3, MLP_H, a1, MLP_H);
        __syncthreads();
        dense_rows<false, RT>(a1, MLP_H, MLP_H, P.wb + MLP_W4, P.wf + MLP_B4, MLP_OUT, o + r0 * MLP_OUT, MLP_OUT);
        __syncthreads();
        if (l < 16 * RT) {
            const uint32_t s = t0 + r0 + l;
            if (s < n) {
                const float *q = o + (r0 + l) * MLP_OUT;
                for (int h = 0; h < 2; ++h) {
                    const V3 v{q[3 * h], q[3 * h + 1], q[3 * h + 2]};
                    const float ln = len(v);
                    out[2 * (size_t)s + h] = ln > 1e-12f ? f4(vdiv(v, ln), 0.0f) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                }
            }
        }
        __syncthreads();
    }
}

// d(1 - cos(p, t))/dp for unit t: -(t / |p| - (p . t) p / |p|^3); zero when |p| ~ 0.
__device__ __forceinline__ float cos_loss(V3 p, V3 t, V3 &g) {
    const float n2 = len2(p);
    if (n2 < 1e-24f) { g = V3{0.0f, 0.0f, 0.0f}; return 1.0f; }
    const float inv = 1.0f / sqrtf(n2);
    const float c = dot(p, t) * inv;
    g = (p * (c * inv) - t) * inv;
    return 1.0f - c;
}

// Gradients of the batch-mean loss over n samples: features BF16 [n][32], targets
// float4 [n][2] (t_gjk, valid) (t_epa, valid).  grad (FP32, NPARAM) and loss (FP32) are
// accumulated with atomics; scale = 1 / batch size.
__global__ void __launch_bounds__(256) k_mlp_grad(const bf16 *feat, const float4 *tgt, uint32_t n, MlpParams P, float scale,
                                                  float *grad, float *loss) {
    __shared__ __attribute__((aligned(16))) bf16 x[MLP_TILE * MLP_IN], a1[MLP_TILE * MLP_H], a2[MLP_TILE * MLP_H],
        a3[MLP_TILE * MLP_H];
    __shared__ __attribute__((aligned(16))) bf16 dz[MLP_TILE * MLP_H], dzp[MLP_TILE * MLP_H], dout[MLP_TILE * MLP_OUT];
    __shared__ float o[MLP_TILE * MLP_OUT];
    const int wave = threadIdx.x >> 6, l = lane64();
    const int r0 = wave * 16;
    float lsum = 0.0f;
    for (uint32_t t0 = blockIdx.x * MLP_TILE; t0 < n; t0 += gridDim.x * MLP_TILE) {
        for (int i = threadIdx.x; i < MLP_TILE * MLP_IN; i += blockDim.x) {
            const uint32_t s = t0 + i / MLP_IN;
            x[i] = s < n ? feat[(size_t)t0 * MLP_IN + i] : (bf16)0.0f;
        }
        __syncthreads();
        mlp_forward16(P, x + r0 * MLP_IN, a1 + r0 * MLP_H, a2 + r0 * MLP_H, a3 + r0 * MLP_H, o + r0 * MLP_OUT);
        // output gradient (one lane per sample)
        if (l < 16) {
            const int r = r0 + l;
            const uint32_t s = t0 + r;
            float d[MLP_OUT];
            for (int k = 0; k < MLP_OUT; ++k) d[k] = 0.0f;
            if (s < n) {
                const float *q = o + r * MLP_OUT;
                for (int h = 0; h < 2; ++h) {
                    const float4 t = tgt[2 * (size_t)s + h];
                    if (t.w != 0.0f) {
                        V3 g;
                        lsum += cos_loss(V3{q[3 * h], q[3 * h + 1], q[3 * h + 2]}, V3{t.x, t.y, t.z}, g) * scale;
                        d[3 * h] = g.x * scale; d[3 * h + 1] = g.y * scale; d[3 * h + 2] = g.z * scale;
                    }
                }
            }
            for (int k = 0; k < MLP_OUT; ++k) dout[r * MLP_OUT + k] = (bf16)d[k];
        }
        __syncthreads();
        // layer 4: dW4 = dout^T a3, db4, dz(a3) = dout W4 masked by a3 > 0
        const bf16 *ins[4] = {x, a1, a2, a3};
        const int kin[4] = {MLP_IN, MLP_H, MLP_H, MLP_H};
        const int nout[4] = {MLP_H, MLP_H, MLP_H, MLP_OUT};
        const int woff[4] = {MLP_W1, MLP_W2, MLP_W3, MLP_W4};
        const int boff[4] = {MLP_B1, MLP_B2, MLP_B3, MLP_B4};
        const bf16 *dcur = dout;
        int ldd = MLP_OUT;
#pragma unroll
        for (int layer = 3; layer >= 0; --layer) {
            const int K = kin[layer], N = nout[layer];
            const bf16 *A = ins[layer];
            // weight gradient: blocks of 16 (n) x 16 (k), contraction over the 64 samples
            const int nbk = (N / 16) * (K / 16);
            for (int blk = wave; blk < nbk; blk += 4) {
                const int nb = (blk / (K / 16)) * 16, kb = (blk % (K / 16)) * 16;
                floatx4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
                acc = mfma16(dcur + nb, 1, ldd, A + kb, K, 1, MLP_TILE, acc);
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    atomicAdd(&grad[woff[layer] + (nb + 4 * (l >> 4) + j) * K + kb + (l & 15)], acc[j]);
            }
            // bias gradient
            for (int c = threadIdx.x; c < N; c += blockDim.x) {
                float sb = 0.0f;
                for (int r = 0; r < MLP_TILE; ++r) sb += bf2f(dcur[r * ldd + c]);
                atomicAdd(&grad[boff[layer] + c], sb);
            }
            if (layer > 0) {
                // input gradient of this wave's rows: dA(s, k) = sum_n dz(s, n) W(n, k), then the
                // ReLU mask of the layer below (its output > 0)
                bf16 *dn = (dcur == dz) ? dzp : dz;
                const bf16 *W = P.wb + woff[layer];
                for (int kb = 0; kb < K; kb += 16) {
                    floatx4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
                    acc = mfma16(dcur + r0 * ldd, ldd, 1, W + kb, K, 1, N, acc);
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int r = r0 + 4 * (l >> 4) + j, k = kb + (l & 15);
                        dn[r * MLP_H + k] = bf2f(A[r * K + k]) > 0.0f ? (bf16)acc[j] : (bf16)0.0f;
                    }
                }
                __syncthreads();
                dcur = dn;
                ldd = MLP_H;
            } else {
                __syncthreads();
            }
        }
    }
    for (int o2 = 32; o2 > 0; o2 >>= 1) lsum += __shfl_xor(lsum, o2, 64);
    if (l == 0 && lsum != 0.0f) atomicAdd(loss, lsum);
}

// Adam step on the FP32 master parameters; refreshes the BF16 copy and clears the gradient.
__global__ void k_mlp_adam(float *wf, bf16 *wb, float *grad, float *m, float *v, uint32_t np, float lr, float b1c, float b2c) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const float g = grad[i];
    const float mi = 0.9f * m[i] + 0.1f * g;
    const float vi = 0.999f * v[i] + 0.001f * g * g;
    m[i] = mi;
    v[i] = vi;
    const float w = wf[i] - lr * (mi / b1c) / (sqrtf(vi / b2c) + 1e-8f);
    wf[i] = w;
    wb[i] = (bf16)w;
    grad[i] = 0.0f;
}

// Training batch: samples (start + i) * mult mod n of the log, i < count.
__global__ void k_mlp_gather(const bf16 *feat, const float4 *tgt, uint64_t n, uint64_t start, uint64_t mult,
                             uint32_t count, bf16 *bfeat, float4 *btgt) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint64_t j = (unsigned __int128)((start + i) % n) * mult % n;
    for (int k = 0; k < MLP_IN; ++k) bfeat[(size_t)i * MLP_IN + k] = feat[j * MLP_IN + k];
    btgt[2 * (size_t)i] = tgt[2 * j];
    btgt[2 * (size_t)i + 1] = tgt[2 * j + 1];
}

// Warm starts given per input pair (slot) into bin order.
__global__ void k_gather_guess(const uint32_t *bin_slot, uint32_t n, const float4 *by_slot, float4 *by_bin) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t s = bin_slot[i];
    by_bin[2 * i] = by_slot[2 * s];
    by_bin[2 * i + 1] = by_slot[2 * s + 1];
}

__global__ void k_bf16_to_f32(const bf16 *in, float *out, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (float)in[i];
}

__global__ void k_f32_to_bf16(const float *in, bf16 *out, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (bf16)in[i];
}

// Shape descriptor: centre and half extents of the local bounds.
__device__ __forceinline__ void shape_descriptor(const Colliders &C, const HullSet &hs, uint32_t c, float *d) {
    const int kind = C.meta[c].x;
    const float4 dm = C.dims[c];
    V3 cen{0.0f, 0.0f, 0.0f}, ext{0.0f, 0.0f, 0.0f};
    if (kind == SH_HULL) {
        const float4 a = hs.amin[C.hull[c]], b = hs.amax[C.hull[c]];
        cen = (V3{a.x, a.y, a.z} + V3{b.x, b.y, b.z}) * 0.5f;
        ext = (V3{b.x, b.y, b.z} - V3{a.x, a.y, a.z}) * 0.5f;
    } else if (kind == SH_SPHERE) {
        ext = V3{dm.x, dm.x, dm.x};
    } else if (kind == SH_BOX) {
        ext = V3{dm.x, dm.y, dm.z};
    } else if (kind == SH_CAPSULE) {
        ext = V3{dm.y, dm.x + dm.y, dm.y};
    }
    d[0] = cen.x; d[1] = cen.y; d[2] = cen.z; d[3] = ext.x; d[4] = ext.y; d[5] = ext.z;
}
__device__ __forceinline__ int kind_slot(int kind) {
    return kind == SH_SPHERE ? 0 : kind == SH_BOX ? 1 : kind == SH_CAPSULE ? 2 : 3;
}

// Features of the pairs bin_slot[off .. off + n) into feat [n][32] (BF16).
__global__ void k_gjk_features(const uint32_t *bin_slot, uint32_t off, uint32_t n, const uint32_t *cand,
                               const uint32_t *ncand, Colliders C, Bodies B, Contacts K, HullSet hs, bf16 *feat,
                               const uint32_t *row0) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;   // candidate (bin position without a list)
    if (i >= gjk_count(n, ncand)) return;
    const uint32_t s = bin_slot[off + (cand ? cand[i] : i)];
    if (row0) feat += (size_t)*row0 * MLP_IN;   // appended after the rows already logged
    const uint2 cc = K.cc[s], bb = K.bb[s];
    const TF ta = collider_tf(C, B, cc.x, bb.x), tb = collider_tf(C, B, cc.y, bb.y);
    float f[MLP_IN];
    const V3 p = tf_inv_vec(ta, tb.pos - ta.pos);
    const M3 R = mat3_cast(qmul(conj(ta.rot), tb.rot));
    f[0] = p.x; f[1] = p.y; f[2] = p.z;
    f[3] = R.c0.x; f[4] = R.c0.y; f[5] = R.c0.z; f[6] = R.c1.x; f[7] = R.c1.y; f[8] = R.c1.z;
    f[9] = R.c2.x; f[10] = R.c2.y; f[11] = R.c2.z;
    shape_descriptor(C, hs, cc.x, f + 12);
    shape_descriptor(C, hs, cc.y, f + 18);
    for (int k = 24; k < 32; ++k) f[k] = 0.0f;
    f[24 + kind_slot(C.meta[cc.x].x)] = 1.0f;
    f[28 + kind_slot(C.meta[cc.y].x)] = 1.0f;
    for (int k = 0; k < MLP_IN; ++k) feat[(size_t)i * MLP_IN + k] = (bf16)f[k];
}

}  // namespace rb
