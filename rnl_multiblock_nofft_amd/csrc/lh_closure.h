// lh_closure.h -- learned closure of the nonlinear term: layouts shared by
// the kernels (lh_closure.hip) and the host entry points (lh_api.hip).
//
// The network: 135 inputs -> 128 -> 128 -> 128 (GELU) -> 3 outputs, bf16
// weights and activations, fp32 accumulation (v_mfma_f32_16x16x32_bf16).
//
// Training layout ("padded"): fp32 master parameters
//   W1[128][160] b1[128] W2[128][128] b2[128] W3[128][128] b3[128]
//   W4[32][128] b4[32]
// (input columns 135..159 and output rows 3..31 stay zero), and bf16 copies
// of the four matrices at the same row/column shapes.
//
// Inference layout ("fragments"): for each layer, output block b (16 rows)
// and k-step s (32 inputs), 64 lanes x 8 bf16 -- exactly the A operand each
// lane of a wave feeds the matrix instruction, so one 16-byte LDS read per
// lane.  Lane l holds row 16b + l%16; its eight k slots are, for layer 1,
// inputs 32s + 8(l/16) + v and, for layers 2-4, the inputs whose values the
// previous layer's accumulators leave in that lane: 32s + 4(l/16) + v for
// v < 4 and 32s + 16 + 4(l/16) + v-4 for v >= 4.  With that order the
// hidden activations never leave registers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lh {

constexpr int CL_NIN = 135, CL_KIN = 160, CL_NH = 128, CL_NOUT = 3, CL_MOUT = 32;
// padded fp32 parameter offsets (floats)
constexpr int CLP_W1 = 0, CLP_B1 = CLP_W1 + CL_NH * CL_KIN, CLP_W2 = CLP_B1 + CL_NH, CLP_B2 = CLP_W2 + CL_NH * CL_NH,
              CLP_W3 = CLP_B2 + CL_NH, CLP_B3 = CLP_W3 + CL_NH * CL_NH, CLP_W4 = CLP_B3 + CL_NH,
              CLP_B4 = CLP_W4 + CL_MOUT * CL_NH, CLP_N = CLP_B4 + CL_MOUT;
// bf16 matrix offsets (elements)
constexpr int CLB_W1 = 0, CLB_W2 = CLB_W1 + CL_NH * CL_KIN, CLB_W3 = CLB_W2 + CL_NH * CL_NH,
              CLB_W4 = CLB_W3 + CL_NH * CL_NH, CLB_N = CLB_W4 + CL_MOUT * CL_NH;
// inference fragments: layer 1 8 x 5, layers 2, 3 8 x 4, layer 4 1 x 4
constexpr int CLF_L1 = 0, CLF_L2 = 40, CLF_L3 = 72, CLF_L4 = 104, CLF_N = 108;
constexpr int CL_FRAG_BYTES = CLF_N * 64 * 16;
constexpr int CL_NBIAS = 3 * CL_NH + 16;  // b1, b2, b3, b4 (padded to 16)
constexpr int CL_MAXPLANE = 64 * 64;      // N1*N2 limit of a closure channel (LDS transforms)

// normalisation: inputs (u - mean[f]) * inv_scale[f], outputs y * out_scale[o]
struct ClNorm {
  float mean[3], inv_scale[3], out_scale[3];
};

// three source arrays (rows [nrows][ldc] of double2) for the plane transforms
struct ClSrc {
  const double2 *f[3];
  int nrows[3];
};

struct Chan;
struct ShapeMax;

// physical N1 x N2 planes (float [3][nyu][N2][N1]) of three mode-major fields
hipError_t launch_cl_phys(char *arena, const Chan *d_chan, const Chan &h, ClSrc src, float *phys, hipStream_t s);
// the network on every point of u rows 1..ny-1: corr [3][nyu][N2][N1]
hipError_t launch_cl_infer(const void *frags, const float *bias, ClNorm nm, const float *phys, float *corr, int N1,
                           int N2, int nyu, int nyv, hipStream_t s);
// forward transform of the correction planes, added to nl1, nl2, nl3
hipError_t launch_cl_to_modes(char *arena, const Chan *d_chan, const Chan &h, const float *corr, hipStream_t s,
                              double2 *store = nullptr);
// add the correction modes a previous k_cl_to_modes stored to nl1, nl2, nl3
hipError_t launch_cl_add_modes(char *arena, const Chan *d_chan, const Chan &h, const double2 *store, hipStream_t s);
// fine -> coarse box-filtered restriction of u1, u2, u3, p (coarse state)
hipError_t launch_cl_restrict_fields(char *fa, const Chan *d_fine, char *ca, const Chan *d_coarse, const Chan &hc,
                                     hipStream_t s);
// box-filtered restriction of the fine channel's full nonlinear term into
// tau [3][nyu][ldc_coarse] (mode = 0) or tau -= the coarse channel's own
// full nonlinear term with the Nyquist modes cleared (mode = 1; fa = ca)
hipError_t launch_cl_tau(char *fa, const Chan *d_src, char *ca, const Chan *d_coarse, const Chan &hc, double2 *tau,
                         int mode, hipStream_t s);

// generic bf16 GEMM with fp32 accumulation: C(m, n) (+)= sum_k A(m, k) B(k, n)
// A(m, k) = A[m*sam + k*sak], B(k, n) = B[k*sbk + n*sbn] (one stride of each
// operand is 1), C(m, n) = C[m*scm + n*scn]; kchunk splits K over blocks
// (then accum must be 1 and C is added to atomically)
struct GemmP {
  const uint16_t *A, *B;
  float *C;
  int M, N, K;
  long sam, sak, sbk, sbn, scm, scn;
  int kchunk, accum;
};
hipError_t launch_gemm(const GemmP &p, hipStream_t s);
// Z[r][n] += b[r]; H = bf16(gelu(Z))
hipError_t launch_bias_gelu(float *Z, const float *b, uint16_t *H, int R, int B, hipStream_t s);
// outputs: d = Y + b4 - T (rows < 3), loss += sum d^2, dY = bf16(2 d scale)
// (rows 3..31 of dY zero); Y [32][B], T [3][B]
hipError_t launch_out_loss(const float *Y, const float *b4, const float *T, uint16_t *dY, double *loss, int B,
                           float scale, hipStream_t s);
// dZ = bf16(dH * gelu'(Z))
hipError_t launch_dgelu(const float *dH, const float *Z, uint16_t *dZ, int R, int B, hipStream_t s);
// db[r] += sum_n D[r][n] (bf16 D)
hipError_t launch_rowsum(const uint16_t *D, float *db, int R, int B, hipStream_t s);
// AdamW over n parameters
hipError_t launch_adam(float *w, const float *g, float *m, float *v, int n, float lr, float b1, float b2, float eps,
                       float wd, float bc1, float bc2, hipStream_t s);
// bf16 copies of the four master matrices
hipError_t launch_pack_bf16(const float *P, uint16_t *Wb, hipStream_t s);
// training batch: B points drawn from nsnap snapshots (splitmix64 of seed and
// sample index), normalised inputs X [160][B] bf16 and targets T [3][B]
hipError_t launch_cl_batch(const float *phys, const float *tau, int nsnap, int N1, int N2, int nyu, int nyv,
                           ClNorm nm, uint64_t seed, uint16_t *X, float *T, int B, hipStream_t s);
// normalised inputs of given host points (tests): X [160][B] from one
// snapshot at points pts[n] (index into u rows 1..ny-1)
hipError_t launch_cl_features(const float *phys, int N1, int N2, int nyu, int nyv, ClNorm nm, const int *pts,
                              uint16_t *X, int B, hipStream_t s);

}  // namespace lh
