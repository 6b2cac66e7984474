/*
 * orc_mpnn.c -- FP64 oracle of the learned initial guess for the pressure
 * solve: a message-passing network on the matrix graph.  Test
 * infrastructure only; see orc.h.  The reference has no such component
 * (it is an addition of this project); this file is the definition the HIP
 * kernels are checked against, and its analytic gradient is pinned by
 * finite differences in tests/test_oracle.py.
 *
 * Network (F = 5 inputs, H hidden, L message-passing layers):
 *   h0      = relu(W_in f + b_in)
 *   m_l     = sum_{m in N(n)} w_nm h_{l-1,m}
 *   h_l     = relu(W_self,l h_{l-1} + W_nbr,l m_l + b_l),   l = 1..L
 *   a_k,n   = w_out,k . h_L + b_out,k,   k = 1, 2
 *   delta_n = a_1,n f_3,n + a_2,n f_4,n   (f_3, f_4 the history features)
 * Edge weights w_nm = sqrt(|a_nm a_mn|) / sqrt(|a_nn a_mm|) (symmetric).
 * History: h_0 = x0 (the newest guess, the previous step's solution),
 * h_1 .. h_{nh-1} the guesses of the solves before (nh <= 5).  Base guess
 * xb = the polynomial extrapolation in time of the newest min(nh, 3)
 * (h_0; 2 h_0 - h_1; 3 h_0 - 3 h_1 + h_2).  Node features: r/s_r
 * (r = b - A xb), log|a_nn| - mean, Dirichlet flag, and the third and
 * fourth backward differences of the history, f_3 = (h_0 - 3 h_1 + 3 h_2
 * - h_3)/s_d (0 when nh < 4), f_4 = (h_0 - 4 h_1 + 6 h_2 - 4 h_3 + h_4)/s_d
 * (0 when nh < 5), with s_d = rms(h_0 - h_1): differences only, no absolute
 * levels, so that what is learned on earlier steps carries over to later
 * ones.  Prediction x = xb + s_o delta with s_o = s_d, or s_o = s_e =
 * rms(r / diag) when h_0 == h_1 (no history); rows without off-diagonal
 * entries take b / a_nn.
 * The network predicts per-node coefficients of the two higher
 * differences: a = (1, 0) turns the quadratic extrapolation into the
 * cubic one, a = (1, 1) into the quartic one, and other values into
 * damped or per-node blends (a learned multi-step extrapolation in time),
 * so its output vanishes with the history and carries no offset of its
 * own.  The untrained head (w_out = 0, b_out = 0) is the quadratic
 * extrapolation.
 * Parameters, flat: W_in[H][F], b_in[H], then per layer W_self[H][H],
 * W_nbr[H][H], b[H], then w_out[2][H], b_out[2].
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "orc.h"

#define NF 5

int64_t orc_mpnn_nparams(int H, int L)
{
  return (int64_t)H * NF + H + (int64_t)L * (2 * H * H + H) + 2 * H + 2;
}

static double rms(const double* v, int64_t n)
{
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i) s += v[i] * v[i];
  s = sqrt(s / (double)(n > 0 ? n : 1));
  return s > 0 ? s : 1.0;
}

static double entry(const orc_csr* A, int64_t r, int32_t c)
{
  for (int64_t k = A->row_ptr[r]; k < A->row_ptr[r + 1]; ++k)
    if (A->cols[k] == c) return A->vals[k];
  return 0.0;
}

void orc_mpnn_slot_weights(const orc_csr* A, double* w)
{
  for (int64_t r = 0; r < A->nrows; ++r) {
    const double dr = fabs(entry(A, r, (int32_t)r));
    for (int64_t k = A->row_ptr[r]; k < A->row_ptr[r + 1]; ++k) {
      const int32_t c = A->cols[k];
      if (c == r) {
        w[k] = 0.0;
        continue;
      }
      const double dc = fabs(entry(A, c, c));
      const double amn = fabs(A->vals[k]) * fabs(entry(A, c, (int32_t)r));
      w[k] = (dr > 0 && dc > 0) ? sqrt(amn) / sqrt(dr * dc) : 0.0;
    }
  }
}

void orc_mpnn_base_guess(int64_t n, const double* const* hist, int nh,
                         double* xb)
{
  for (int64_t i = 0; i < n; ++i) {
    const double h0 = hist[0][i];
    xb[i] = nh >= 3 ? 3.0 * h0 - 3.0 * hist[1][i] + hist[2][i]
          : nh == 2 ? 2.0 * h0 - hist[1][i]
                    : h0;
  }
}

void orc_mpnn_features(const orc_csr* A, const double* b,
                       const double* const* hist, int nh, double* feat,
                       double* scales)
{
  const int64_t n = A->nrows;
  double* xb = (double*)malloc(sizeof(double) * (size_t)n);
  double* r = (double*)malloc(sizeof(double) * (size_t)n);
  double* e = (double*)malloc(sizeof(double) * (size_t)n);
  double* d = (double*)malloc(sizeof(double) * (size_t)n);
  const double* h0 = hist[0];
  orc_mpnn_base_guess(n, hist, nh, xb);
  orc_spmv(A, xb, r);
  double mlog = 0.0, dd = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    r[i] = b[i] - r[i];
    const double a = fabs(entry(A, i, (int32_t)i));
    e[i] = a > 0 ? r[i] / a : 0.0;
    d[i] = nh >= 2 ? h0[i] - hist[1][i] : 0.0;
    dd += d[i] * d[i];
    mlog += log(a > 0 ? a : 1.0);
  }
  mlog /= (double)(n > 0 ? n : 1);
  const double sb = rms(b, n), sx = rms(h0, n), sr = rms(r, n), se = rms(e, n);
  const double sd = rms(d, n);
  for (int64_t i = 0; i < n; ++i) {
    const double a = fabs(entry(A, i, (int32_t)i));
    int dir = 1;
    for (int64_t k = A->row_ptr[i]; k < A->row_ptr[i + 1]; ++k)
      if (A->cols[k] != i && A->vals[k] != 0.0) dir = 0;
    feat[NF * i + 0] = r[i] / sr;
    feat[NF * i + 1] = log(a > 0 ? a : 1.0) - mlog;
    feat[NF * i + 2] = (double)dir;
    feat[NF * i + 3] =
      nh >= 4 ? (h0[i] - 3.0 * hist[1][i] + 3.0 * hist[2][i] - hist[3][i]) / sd
              : 0.0;
    feat[NF * i + 4] =
      nh >= 5 ? (h0[i] - 4.0 * hist[1][i] + 6.0 * hist[2][i] -
                 4.0 * hist[3][i] + hist[4][i]) / sd
              : 0.0;
  }
  scales[0] = sb;
  scales[1] = sx;
  scales[2] = sr;
  scales[3] = se;
  scales[4] = dd > 0 ? sd : se; /* output scale */
  free(xb);
  free(r);
  free(e);
  free(d);
}

static void aggregate(const orc_csr* A, const double* w, int H,
                      const double* h, double* m)
{
  for (int64_t i = 0; i < A->nrows; ++i) {
    for (int f = 0; f < H; ++f) m[H * i + f] = 0.0;
    for (int64_t k = A->row_ptr[i]; k < A->row_ptr[i + 1]; ++k) {
      const double wk = w[k];
      if (wk == 0.0) continue;
      const int32_t c = A->cols[k];
      for (int f = 0; f < H; ++f) m[H * i + f] += wk * h[(int64_t)H * c + f];
    }
  }
}

/* Forward; if acts != NULL it receives z_0..z_L and m_1..m_L for the
 * backward: acts = [z0 (N*H)] [m_l, z_l for l = 1..L] */
void orc_mpnn_forward(const double* P, int H, int L, const orc_csr* A,
                      const double* w, const double* feat, double* delta,
                      double* acts)
{
  const int64_t n = A->nrows;
  const double* Win = P;
  const double* bin = Win + (int64_t)H * NF;
  double* h = (double*)malloc(sizeof(double) * (size_t)(n * H));
  double* h2 = (double*)malloc(sizeof(double) * (size_t)(n * H));
  double* m = (double*)malloc(sizeof(double) * (size_t)(n * H));
  for (int64_t i = 0; i < n; ++i)
    for (int f = 0; f < H; ++f) {
      double z = bin[f];
      for (int k = 0; k < NF; ++k) z += Win[f * NF + k] * feat[NF * i + k];
      if (acts) acts[H * i + f] = z;
      h[H * i + f] = z > 0 ? z : 0.0;
    }
  const double* Pl = bin + H;
  for (int l = 0; l < L; ++l) {
    const double* Ws = Pl;
    const double* Wn = Ws + H * H;
    const double* bl = Wn + H * H;
    aggregate(A, w, H, h, m);
    double* ml = acts ? acts + n * H * (1 + 2 * l) : NULL;
    double* zl = acts ? acts + n * H * (2 + 2 * l) : NULL;
    if (ml) memcpy(ml, m, sizeof(double) * (size_t)(n * H));
    for (int64_t i = 0; i < n; ++i)
      for (int f = 0; f < H; ++f) {
        double z = bl[f];
        for (int k = 0; k < H; ++k)
          z += Ws[f * H + k] * h[H * i + k] + Wn[f * H + k] * m[H * i + k];
        if (zl) zl[H * i + f] = z;
        h2[H * i + f] = z > 0 ? z : 0.0;
      }
    double* t = h;
    h = h2;
    h2 = t;
    Pl = bl + H;
  }
  const double* wout = Pl; /* [2][H], then b_out[2] */
  for (int64_t i = 0; i < n; ++i) {
    double a1 = wout[2 * H], a2 = wout[2 * H + 1];
    for (int f = 0; f < H; ++f) {
      a1 += wout[f] * h[H * i + f];
      a2 += wout[H + f] * h[H * i + f];
    }
    delta[i] = a1 * feat[NF * i + 3] + a2 * feat[NF * i + 4];
  }
  free(h);
  free(h2);
  free(m);
}

/* Nearest-even bfloat16 rounding of x taken to FP32 first (the device
 * rounds FP32 values). */
static double bf16_round(double x)
{
  float f = (float)x;
  uint32_t u;
  memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  u &= 0xFFFF0000u;
  memcpy(&f, &u, 4);
  return (double)f;
}

static double fl(double x) { return (double)(float)x; }

/* The reduced-precision forward of the prediction path (NW_MPNN_BF16):
 * the hidden activations h_0..h_{L-1}, the messages m_l and the layer
 * weights W_self, W_nbr are rounded to bfloat16; the features, slot
 * weights, W_in, the biases and the output head are rounded to FP32; the
 * last layer's activations stay unrounded; sums in FP64. */
void orc_mpnn_forward_bf16(const double* P, int H, int L, const orc_csr* A,
                           const double* w, const double* feat, double* delta)
{
  const int64_t n = A->nrows;
  const double* Win = P;
  const double* bin = Win + (int64_t)H * NF;
  double* h = (double*)malloc(sizeof(double) * (size_t)(n * H));
  double* h2 = (double*)malloc(sizeof(double) * (size_t)(n * H));
  double* m = (double*)malloc(sizeof(double) * (size_t)(n * H));
  double* wf = (double*)malloc(sizeof(double) * (size_t)A->row_ptr[n]);
  for (int64_t k = 0; k < A->row_ptr[n]; ++k) wf[k] = fl(w[k]);
  for (int64_t i = 0; i < n; ++i)
    for (int f = 0; f < H; ++f) {
      double z = fl(bin[f]);
      for (int k = 0; k < NF; ++k) z += fl(Win[f * NF + k]) * fl(feat[NF * i + k]);
      h[H * i + f] = bf16_round(z > 0 ? z : 0.0);
    }
  const double* Pl = bin + H;
  for (int l = 0; l < L; ++l) {
    const double* Ws = Pl;
    const double* Wn = Ws + H * H;
    const double* bl = Wn + H * H;
    aggregate(A, wf, H, h, m);
    for (int64_t i = 0; i < n * H; ++i) m[i] = bf16_round(m[i]);
    for (int64_t i = 0; i < n; ++i)
      for (int f = 0; f < H; ++f) {
        double z = fl(bl[f]);
        for (int k = 0; k < H; ++k)
          z += bf16_round(Ws[f * H + k]) * h[H * i + k] +
               bf16_round(Wn[f * H + k]) * m[H * i + k];
        const double r = z > 0 ? z : 0.0;
        h2[H * i + f] = l + 1 < L ? bf16_round(r) : r;
      }
    double* t = h;
    h = h2;
    h2 = t;
    Pl = bl + H;
  }
  const double* wout = Pl;
  for (int64_t i = 0; i < n; ++i) {
    double a1 = fl(wout[2 * H]), a2 = fl(wout[2 * H + 1]);
    for (int f = 0; f < H; ++f) {
      a1 += fl(wout[f]) * h[H * i + f];
      a2 += fl(wout[H + f]) * h[H * i + f];
    }
    delta[i] = a1 * fl(feat[NF * i + 3]) + a2 * fl(feat[NF * i + 4]);
  }
  free(h);
  free(h2);
  free(m);
  free(wf);
}

/* Backward of the network from dd = dLoss/ddelta (n values); acts from
 * orc_mpnn_forward; grad (nparams) is overwritten with dLoss/dP. */
static void mpnn_backward(const double* P, int H, int L, const orc_csr* A,
                          const double* w, const double* feat,
                          const double* acts, const double* dd, double* grad)
{
  const int64_t n = A->nrows;
  memset(grad, 0, sizeof(double) * (size_t)orc_mpnn_nparams(H, L));
  /* locate parameter blocks */
  const double* Win = P;
  const double* bin = Win + (int64_t)H * NF;
  double* gWin = grad;
  double* gbin = gWin + (int64_t)H * NF;
  const double* wout = bin + H + (int64_t)L * (2 * H * H + H);
  double* gwout = gbin + H + (int64_t)L * (2 * H * H + H);
  /* h_L = relu(z_L) */
  double* g = (double*)malloc(sizeof(double) * (size_t)(n * H));   /* dL/dh */
  double* gz = (double*)malloc(sizeof(double) * (size_t)(n * H));
  double* q = (double*)malloc(sizeof(double) * (size_t)(n * H));
  double* gm = (double*)malloc(sizeof(double) * (size_t)(n * H));
  const double* zL = acts + n * H * (2 * L);
  for (int64_t i = 0; i < n; ++i) {
    const double g1 = dd[i] * feat[NF * i + 3], g2 = dd[i] * feat[NF * i + 4];
    gwout[2 * H] += g1;
    gwout[2 * H + 1] += g2;
    for (int f = 0; f < H; ++f) {
      const double hf = zL[H * i + f] > 0 ? zL[H * i + f] : 0.0;
      gwout[f] += g1 * hf;
      gwout[H + f] += g2 * hf;
      g[H * i + f] = g1 * wout[f] + g2 * wout[H + f];
    }
  }
  for (int l = L - 1; l >= 0; --l) {
    const double* Ws = bin + H + (int64_t)l * (2 * H * H + H);
    const double* Wn = Ws + H * H;
    double* gWs = gbin + H + (int64_t)l * (2 * H * H + H);
    double* gWn = gWs + H * H;
    double* gbl = gWn + H * H;
    const double* zl = acts + n * H * (2 + 2 * l);
    const double* ml = acts + n * H * (1 + 2 * l);
    const double* zprev = acts + (l == 0 ? 0 : n * H * (2 * l));
    for (int64_t i = 0; i < n * H; ++i) gz[i] = zl[i] > 0 ? g[i] : 0.0;
    for (int64_t i = 0; i < n; ++i)
      for (int f = 0; f < H; ++f) {
        const double gzf = gz[H * i + f];
        gbl[f] += gzf;
        for (int k = 0; k < H; ++k) {
          const double hp = zprev[H * i + k] > 0 ? zprev[H * i + k] : 0.0;
          gWs[f * H + k] += gzf * hp;
          gWn[f * H + k] += gzf * ml[H * i + k];
        }
      }
    /* dL/dh_{l-1} = Ws^T gz + aggregate(Wn^T gz) (weights are symmetric) */
    for (int64_t i = 0; i < n; ++i)
      for (int k = 0; k < H; ++k) {
        double a = 0.0, b = 0.0;
        for (int f = 0; f < H; ++f) {
          a += Ws[f * H + k] * gz[H * i + f];
          b += Wn[f * H + k] * gz[H * i + f];
        }
        g[H * i + k] = a;
        q[H * i + k] = b;
      }
    aggregate(A, w, H, q, gm);
    for (int64_t i = 0; i < n * H; ++i) g[i] += gm[i];
  }
  /* input layer */
  for (int64_t i = 0; i < n; ++i)
    for (int f = 0; f < H; ++f) {
      const double gzf = acts[H * i + f] > 0 ? g[H * i + f] : 0.0;
      gbin[f] += gzf;
      for (int k = 0; k < NF; ++k) gWin[f * NF + k] += gzf * feat[NF * i + k];
    }
  free(g);
  free(gz);
  free(q);
  free(gm);
}

/* Loss = mean((delta - t)^2); grad (nparams) receives dLoss/dP. */
double orc_mpnn_loss_grad(const double* P, int H, int L, const orc_csr* A,
                          const double* w, const double* feat,
                          const double* target, double* grad)
{
  const int64_t n = A->nrows;
  double* acts = (double*)malloc(sizeof(double) * (size_t)(n * H * (1 + 2 * L)));
  double* delta = (double*)malloc(sizeof(double) * (size_t)n);
  double* dd = (double*)malloc(sizeof(double) * (size_t)n);
  orc_mpnn_forward(P, H, L, A, w, feat, delta, acts);
  double loss = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    const double e = delta[i] - target[i];
    loss += e * e;
    dd[i] = 2.0 * e / (double)n;
  }
  loss /= (double)n;
  mpnn_backward(P, H, L, A, w, feat, acts, dd, grad);
  free(acts);
  free(delta);
  free(dd);
  return loss;
}

/* Residual loss: u = delta - t on the rows with couplings (0 on Dirichlet
 * rows, feature 2 set), v = D^-1 A u, Loss = mean(v^2): the Jacobi-scaled
 * residual of the prediction error (in units of the output scale), the
 * quantity GMRES starts from.  dLoss/du = (2/n) A^T D^-1 v, masked like u. */
double orc_mpnn_loss_grad_residual(const double* P, int H, int L,
                                   const orc_csr* A, const double* w,
                                   const double* feat, const double* target,
                                   double* grad)
{
  const int64_t n = A->nrows;
  double* acts = (double*)malloc(sizeof(double) * (size_t)(n * H * (1 + 2 * L)));
  double* delta = (double*)malloc(sizeof(double) * (size_t)n);
  double* u = (double*)malloc(sizeof(double) * (size_t)n);
  double* vd = (double*)malloc(sizeof(double) * (size_t)n);
  double* dd = (double*)malloc(sizeof(double) * (size_t)n);
  orc_mpnn_forward(P, H, L, A, w, feat, delta, acts);
  for (int64_t i = 0; i < n; ++i)
    u[i] = feat[NF * i + 2] != 0.0 ? 0.0 : delta[i] - target[i];
  double loss = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    double au = 0.0, di = 0.0;
    for (int64_t k = A->row_ptr[i]; k < A->row_ptr[i + 1]; ++k) {
      au += A->vals[k] * u[A->cols[k]];
      if (A->cols[k] == i) di = A->vals[k];
    }
    const double v = au / di;
    loss += v * v;
    vd[i] = v / di;  /* D^-1 v */
  }
  loss /= (double)n;
  for (int64_t i = 0; i < n; ++i) dd[i] = 0.0;
  for (int64_t i = 0; i < n; ++i)  /* dd = (2/n) A^T vd */
    for (int64_t k = A->row_ptr[i]; k < A->row_ptr[i + 1]; ++k)
      dd[A->cols[k]] += 2.0 / (double)n * A->vals[k] * vd[i];
  for (int64_t i = 0; i < n; ++i)
    if (feat[NF * i + 2] != 0.0) dd[i] = 0.0;
  mpnn_backward(P, H, L, A, w, feat, acts, dd, grad);
  free(acts);
  free(delta);
  free(u);
  free(vd);
  free(dd);
  return loss;
}
