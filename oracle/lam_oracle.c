/*
 * lam_oracle.c -- CPU restatement of the LAMBDA correlate -> beamform ->
 * RFI-null hot path (jdgsmallwood/cuda-spatial-filtering).
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for the HIP product path and
 * the timed CPU baseline.  The product never links or calls it.
 *
 * Exact integer arithmetic for ingest and correlation, IEEE double for
 * everything else.  See lam_oracle.h for layouts and per-function citations.
 */
#define _GNU_SOURCE
#include "lam_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* Threads of the following oracle calls (the CPU baseline's single-core figure). */
void orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* ------------------------------------------------------------------ ingest */
/* packet_formats.hpp:309-312 gives the ring layout; spatial.cuh:430-486 applies
 * the int16 scale of (channel, ring slot, receiver, pol) to the int8 sample;
 * spatial.cuh:1431-1436 starts the adaptive window at slot 1 shifted by the
 * FPGA's sample delay; spatial.cuh:254-264 reads slots 0..NP-1 for corr-beam;
 * spatial.cuh:610-640 maps each output (receiver, pol) to a source stream or
 * to zero.  The scale is looked up for the sample's own receiver (the
 * index zero_missing_packets writes, packet_formats.hpp:103). */
void orc_ingest(const orc_shape *s, const int8_t *samples, const int16_t *scales,
                int window, const int32_t *delays, const int32_t *recv_perm,
                const int32_t *pol_perm, int32_t *y) {
  const int C = s->C, F = s->F, NT = s->NT, R = s->R, P = s->P, RPP = s->RPP,
            NP = s->NP;
  const long T = (long)NP * NT;
  const long slot_stride = (long)F * NT * RPP * P * 2;
  const long chan_stride = (long)(NP + 2) * slot_stride;
  for (int c = 0; c < C; ++c)
    for (long t = 0; t < T; ++t)
      for (int r = 0; r < R; ++r)
        for (int p = 0; p < P; ++p) {
          int src = r, sp = p;
          int32_t *out = y + ((((long)c * T + t) * R + r) * P + p) * 2;
          if (recv_perm) {
            src = recv_perm[r * P + p];
            sp = pol_perm ? pol_perm[r * P + p] : p;
            if (src < 0) {
              out[0] = out[1] = 0;
              continue;
            }
          }
          const int f = src / RPP, n = src % RPP;
          long g = t;
          if (window == ORC_WINDOW_ADAPTIVE) g += NT + (delays ? delays[f] : 0);
          const long slot = g / NT, tt = g % NT;
          const int8_t *x = samples + (long)c * chan_stride + slot * slot_stride +
                            (((long)f * NT + tt) * RPP + n) * P * 2 + sp * 2;
          const int32_t sc =
              scales[(((long)c * (NP + 2) + slot) * R + src) * P + sp];
          out[0] = (int32_t)x[0] * sc;
          out[1] = (int32_t)x[1] * sc;
        }
}

/* -------------------------------------------------------------- correlator */
void orc_correlate(const orc_shape *s, const int32_t *y, int64_t *vis) {
  const int C = s->C, R = s->R, P = s->P;
  const long T = (long)s->NP * s->NT;
  const long nbl = (long)R * (R + 1) / 2;
#pragma omp parallel for schedule(dynamic)
  for (int c = 0; c < C; ++c) {
    int64_t *vc = vis + (long)c * nbl * P * P * 2;
    memset(vc, 0, sizeof(int64_t) * nbl * P * P * 2);
    const int32_t *yc = y + (long)c * T * R * P * 2;
    for (long t = 0; t < T; ++t) {
      const int32_t *yt = yc + t * R * P * 2;
      for (int j = 0; j < R; ++j)
        for (int i = 0; i <= j; ++i) {
          int64_t *v = vc + ((long)j * (j + 1) / 2 + i) * P * P * 2;
          for (int p = 0; p < P; ++p) {
            const int64_t ar = yt[(i * P + p) * 2], ai = yt[(i * P + p) * 2 + 1];
            for (int q = 0; q < P; ++q) {
              const int64_t br = yt[(j * P + q) * 2],
                            bi = yt[(j * P + q) * 2 + 1];
              /* (ar + i ai) * (br - i bi) */
              v[(p * P + q) * 2] += ar * br + ai * bi;
              v[(p * P + q) * 2 + 1] += ai * br - ar * bi;
            }
          }
        }
    }
  }
}

/* -------------------------------------------------------------- beamformer */
void orc_beamform(const orc_shape *s, const int32_t *y, const double *w, int nb,
                  double *beams) {
  const int C = s->C, R = s->R, P = s->P;
  const long T = (long)s->NP * s->NT;
#pragma omp parallel for collapse(2) schedule(static)
  for (int c = 0; c < C; ++c)
    for (int p = 0; p < P; ++p)
      for (int b = 0; b < nb; ++b) {
        const double *wb = w + (((long)c * P + p) * nb + b) * R * 2;
        double *out = beams + ((((long)c * P + p) * nb + b) * T) * 2;
        for (long t = 0; t < T; ++t) {
          const int32_t *yt = y + ((long)c * T + t) * R * P * 2;
          double re = 0.0, im = 0.0;
          for (int r = 0; r < R; ++r) {
            const double xr = yt[(r * P + p) * 2], xi = yt[(r * P + p) * 2 + 1];
            re += wb[2 * r] * xr - wb[2 * r + 1] * xi;
            im += wb[2 * r] * xi + wb[2 * r + 1] * xr;
          }
          out[2 * t] = re;
          out[2 * t + 1] = im;
        }
      }
}

/* --------------------------------------------------------------- covariance */
void orc_covariance(int R, int P, const int64_t *vis_c, int p, double *A) {
  for (int j = 0; j < R; ++j)
    for (int i = 0; i <= j; ++i) {
      const int64_t *v = vis_c + ((long)j * (j + 1) / 2 + i) * P * P * 2 +
                         (p * P + p) * 2;
      A[2 * ((long)j * R + i)] = (double)v[0];
      A[2 * ((long)j * R + i) + 1] = (double)v[1];
      A[2 * ((long)i * R + j)] = (double)v[0];
      A[2 * ((long)i * R + j) + 1] = -(double)v[1];
    }
}

void orc_whiten(int n, double *A, double *d) {
  for (int k = 0; k < n; ++k) {
    const double dre = A[2 * ((long)k * n + k)];
    d[k] = sqrt(dre > 1e-10 ? dre : 1e-10);
  }
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      const double den = d[i] * d[j];
      A[2 * ((long)j * n + i)] /= den;
      A[2 * ((long)j * n + i) + 1] /= den;
    }
}

/* ------------------------------------------------------------------ Jacobi */
/* Complex Jacobi rotation J = diag(1, e^{-i phi}) [[c, s], [-s, c]] that
 * annihilates a_pq = |b| e^{i phi}: tau = (a_qq - a_pp) / (2|b|),
 * t = sgn(tau) / (|tau| + sqrt(1 + tau^2)), c = 1/sqrt(1+t^2), s = t c.
 * A <- J^H A J and V <- V J. */
int orc_jacobi_eigh(int n, double *A, double *evals, double *V, int max_sweeps,
                    double tol) {
#define AR(i, j) A[2 * ((long)(j) * n + (i))]
#define AI(i, j) A[2 * ((long)(j) * n + (i)) + 1]
#define VR(i, j) V[2 * ((long)(j) * n + (i))]
#define VI(i, j) V[2 * ((long)(j) * n + (i)) + 1]
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      VR(i, j) = (i == j) ? 1.0 : 0.0;
      VI(i, j) = 0.0;
    }
  /* Hermitian: keep the diagonal exactly real. */
  for (int i = 0; i < n; ++i) AI(i, i) = 0.0;
  int sweep = 0;
  for (; sweep < max_sweeps; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        const double m = AR(i, j) * AR(i, j) + AI(i, j) * AI(i, j);
        tot += m;
        if (i != j) off += m;
      }
    if (off == 0.0 || off <= tol * tol * tot) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double br = AR(p, q), bi = AI(p, q);
        const double babs = hypot(br, bi);
        if (babs == 0.0) continue;
        const double a = AR(p, p), d = AR(q, q);
        const double tau = (d - a) / (2.0 * babs);
        const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
        const double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
        const double er = br / babs, ei = bi / babs; /* e^{i phi} */
        /* columns: new_p = c X_p - s e^{-i phi} X_q ; new_q = s X_p + c e^{-i phi} X_q */
        for (int i = 0; i < n; ++i) {
          const double xpr = AR(i, p), xpi = AI(i, p), xqr = AR(i, q), xqi = AI(i, q);
          const double yr = er * xqr + ei * xqi, yi = er * xqi - ei * xqr; /* e^{-i phi} x_q */
          AR(i, p) = c * xpr - s * yr;
          AI(i, p) = c * xpi - s * yi;
          AR(i, q) = s * xpr + c * yr;
          AI(i, q) = s * xpi + c * yi;
        }
        /* rows: new_p = c X_p - s e^{i phi} X_q ; new_q = s X_p + c e^{i phi} X_q */
        for (int j = 0; j < n; ++j) {
          const double xpr = AR(p, j), xpi = AI(p, j), xqr = AR(q, j), xqi = AI(q, j);
          const double yr = er * xqr - ei * xqi, yi = er * xqi + ei * xqr; /* e^{i phi} x_q */
          AR(p, j) = c * xpr - s * yr;
          AI(p, j) = c * xpi - s * yi;
          AR(q, j) = s * xpr + c * yr;
          AI(q, j) = s * xpi + c * yi;
        }
        AR(p, q) = AI(p, q) = AR(q, p) = AI(q, p) = 0.0;
        AI(p, p) = AI(q, q) = 0.0;
        for (int i = 0; i < n; ++i) {
          const double xpr = VR(i, p), xpi = VI(i, p), xqr = VR(i, q), xqi = VI(i, q);
          const double yr = er * xqr + ei * xqi, yi = er * xqi - ei * xqr;
          VR(i, p) = c * xpr - s * yr;
          VI(i, p) = c * xpi - s * yi;
          VR(i, q) = s * xpr + c * yr;
          VI(i, q) = s * xpi + c * yi;
        }
      }
  }
  /* ascending selection sort of eigenpairs */
  for (int i = 0; i < n; ++i) evals[i] = AR(i, i);
  for (int i = 0; i < n - 1; ++i) {
    int m = i;
    for (int j = i + 1; j < n; ++j)
      if (evals[j] < evals[m]) m = j;
    if (m != i) {
      const double tv = evals[i];
      evals[i] = evals[m];
      evals[m] = tv;
      for (int r = 0; r < n; ++r) {
        double tr = VR(r, i), ti = VI(r, i);
        VR(r, i) = VR(r, m);
        VI(r, i) = VI(r, m);
        VR(r, m) = tr;
        VI(r, m) = ti;
      }
    }
  }
  return sweep;
#undef AR
#undef AI
#undef VR
#undef VI
}

/* ----------------------------------------------------------- detect / null */
double orc_percentile_linear(const double *sorted, int n, double q) {
  if (n <= 0) return 0.0;
  if (n == 1) return sorted[0];
  const double qc = q < 0.0 ? 0.0 : (q > 1.0 ? 1.0 : q);
  const double pos = qc * (double)(n - 1);
  const int lo = (int)floor(pos), hi = (int)ceil(pos);
  if (lo == hi) return sorted[lo];
  return sorted[lo] + (pos - lo) * (sorted[hi] - sorted[lo]);
}

int orc_detect_k(const double *ev, int n, double delta) {
  const double p20 = orc_percentile_linear(ev, n, 0.2);
  const double p50 = orc_percentile_linear(ev, n, 0.5);
  const double p80 = orc_percentile_linear(ev, n, 0.8);
  const double thr = p50 + delta * ((p80 - p20) / (2.0 * 0.8416));
  int k = 0;
  for (int i = 0; i < n; ++i)
    if (ev[i] > thr) ++k;
  return k;
}

void orc_unwhiten_normalize(int n, double *V, const double *d) {
  for (int k = 0; k < n; ++k) {
    double nrm = 0.0;
    for (int i = 0; i < n; ++i) {
      V[2 * ((long)k * n + i)] *= d[i];
      V[2 * ((long)k * n + i) + 1] *= d[i];
      nrm += V[2 * ((long)k * n + i)] * V[2 * ((long)k * n + i)] +
             V[2 * ((long)k * n + i) + 1] * V[2 * ((long)k * n + i) + 1];
    }
    const double inv = nrm > 0.0 ? 1.0 / sqrt(nrm) : 0.0;
    for (int i = 0; i < n; ++i) {
      V[2 * ((long)k * n + i)] *= inv;
      V[2 * ((long)k * n + i) + 1] *= inv;
    }
  }
}

static int clamp_k(int k, int n) { return k < 0 ? 0 : (k > n ? n : k); }

void orc_null_projection(int n, const double *V, int k, double *P) {
  k = clamp_k(k, n);
  for (int col = 0; col < n; ++col)
    for (int row = 0; row < n; ++row) {
      double sr = 0.0, si = 0.0;
      for (int m = n - k; m < n; ++m) {
        const double ur = V[2 * ((long)m * n + row)], ui = V[2 * ((long)m * n + row) + 1];
        const double vr = V[2 * ((long)m * n + col)], vi = V[2 * ((long)m * n + col) + 1];
        sr += ur * vr + ui * vi; /* u_row * conj(u_col) */
        si += ui * vr - ur * vi;
      }
      P[2 * ((long)col * n + row)] = (row == col ? 1.0 : 0.0) - sr;
      P[2 * ((long)col * n + row) + 1] = -si;
    }
}

void orc_shrink_projection(int n, const double *V, const double *ev, int k,
                           double *P) {
  k = clamp_k(k, n);
  double *dsc = (double *)malloc(sizeof(double) * n);
  double lbar = 0.0;
  if (k < n) {
    for (int i = 0; i < n - k; ++i) lbar += ev[i];
    lbar /= (double)(n - k);
  }
  for (int i = 0; i < n; ++i)
    dsc[i] = (i < n - k) ? 1.0 : (ev[i] > 0.0 ? lbar / ev[i] : 0.0);
  for (int col = 0; col < n; ++col)
    for (int row = 0; row < n; ++row) {
      double sr = 0.0, si = 0.0;
      for (int m = 0; m < n; ++m) {
        const double ur = V[2 * ((long)m * n + row)], ui = V[2 * ((long)m * n + row) + 1];
        const double vr = V[2 * ((long)m * n + col)], vi = V[2 * ((long)m * n + col) + 1];
        sr += dsc[m] * (ur * vr + ui * vi);
        si += dsc[m] * (ui * vr - ur * vi);
      }
      P[2 * ((long)col * n + row)] = sr;
      P[2 * ((long)col * n + row) + 1] = si;
    }
  free(dsc);
}

void orc_apply_projection(int nb, int n, const double *W, const double *P,
                          double *Wp) {
  for (int b = 0; b < nb; ++b)
    for (int col = 0; col < n; ++col) {
      double sr = 0.0, si = 0.0;
      for (int k = 0; k < n; ++k) {
        const double wr = W[2 * ((long)b * n + k)], wi = W[2 * ((long)b * n + k) + 1];
        const double pr = P[2 * ((long)col * n + k)], pi = P[2 * ((long)col * n + k) + 1];
        sr += wr * pr - wi * pi;
        si += wr * pi + wi * pr;
      }
      Wp[2 * ((long)b * n + col)] = sr;
      Wp[2 * ((long)b * n + col) + 1] = si;
    }
}

void orc_fine_delay_table(int C, int R, int N, const float *delays_ns,
                          double base_freq_hz, double channel_bw_hz,
                          int min_freq_ch, double *table) {
  for (int c = 0; c < C; ++c) {
    const double fc = base_freq_hz + (double)(min_freq_ch + c) * channel_bw_hz;
    for (int r = 0; r < R; ++r) {
      const double tau = (double)delays_ns[r] * 1e-9;
      for (int k = 0; k < N; ++k) {
        const long ks = (k < N / 2) ? k : (long)k - N;
        const double ff = fc + ((double)ks / N) * channel_bw_hz;
        const double phi = -2.0 * M_PI * ff * tau;
        table[2 * (((long)c * R + r) * N + k)] = cos(phi);
        table[2 * (((long)c * R + r) * N + k) + 1] = sin(phi);
      }
    }
  }
}

/* --------------------------------------------------------- beam spectra */
void orc_beam_spectra(int C, int P, int E, int np, int nfft, int drop,
                      const double *beams, double *out) {
  const int keep = nfft / 2 - drop, nout = nfft - 2 * drop;
  const long rows = (long)C * P * E;
  for (long row = 0; row < rows; ++row) {
    const double *x = beams + 2 * row * (long)np * nfft;
    double *o = out + row * nout;
    for (int g = 0; g < nout; ++g) o[g] = 0.0;
    for (int pk = 0; pk < np; ++pk) {
      const double *xp = x + 2L * pk * nfft;
      for (int g = 0; g < nout; ++g) {
        const int u = g < keep ? g + nfft / 2 + drop : g - keep;
        double re = 0.0, im = 0.0;
        for (int t = 0; t < nfft; ++t) {
          /* exact phase index: u*t mod nfft, so the twiddle is one of nfft values */
          const double a = -2.0 * M_PI * (double)((long)u * t % nfft) / nfft;
          const double c = cos(a), s = sin(a);
          re += xp[2 * t] * c - xp[2 * t + 1] * s;
          im += xp[2 * t] * s + xp[2 * t + 1] * c;
        }
        o[g] += sqrt(re * re + im * im);
      }
    }
  }
}

void orc_bandpass(int rows, int n, int d, const double *x, double *out) {
  for (long row = 0; row < rows; ++row) {
    const double *xr = x + 2 * row * (long)n;
    for (int m = 0; m < n / d; ++m) {
      double sum = 0.0;
      int cnt = 0;
      for (int k = m * d; k < (m + 1) * d; ++k) {
        double re = 0.0, im = 0.0;
        for (int t = 0; t < n; ++t) {
          const double a = -2.0 * M_PI * (double)((long)k * t % n) / n;
          const double c = cos(a), s = sin(a);
          re += xr[2 * t] * c - xr[2 * t + 1] * s;
          im += xr[2 * t] * s + xr[2 * t + 1] * c;
        }
        const double p = re * re + im * im;
        if (!isnan(p)) {
          sum += p;
          ++cnt;
        }
      }
      out[row * (n / d) + m] = cnt ? sum / cnt : 0.0;
    }
  }
}

/* ------------------------------------------------------------ block driver */
int orc_rfi_null_cp(int R, int P, int B, const int64_t *vis_c, int p,
                    const double *w_cp, int fixed_k, double delta, int shrink,
                    double *out_w, double *evals, double *evecs) {
  const int n = R;
  double *A = (double *)malloc(sizeof(double) * 2 * n * n);
  double *Pm = (double *)malloc(sizeof(double) * 2 * n * n);
  double *d = (double *)malloc(sizeof(double) * n);
  orc_covariance(R, P, vis_c, p, A);
  orc_whiten(n, A, d);
  orc_jacobi_eigh(n, A, evals, evecs, 100, 1e-15);
  int k = fixed_k >= 0 ? fixed_k : orc_detect_k(evals, n, delta);
  if (!shrink) {
    orc_unwhiten_normalize(n, evecs, d);
    orc_null_projection(n, evecs, k, Pm);
  } else {
    orc_shrink_projection(n, evecs, evals, k, Pm);
  }
  memcpy(out_w, w_cp, sizeof(double) * 2 * B * n);
  orc_apply_projection(B, n, w_cp, Pm, out_w + 2 * (long)B * n);
  free(A);
  free(Pm);
  free(d);
  return k;
}

int orc_adaptive_block(const orc_shape *s, const int32_t *y, const double *w,
                       const int32_t *fixed_k, int detect, double delta,
                       int shrink, int64_t *vis, double *beams2b,
                       double *evals, double *evecs, int32_t *counts) {
  const int C = s->C, R = s->R, P = s->P, B = s->B;
  const long T = (long)s->NP * s->NT;
  const long nbl = (long)R * (R + 1) / 2;
  orc_correlate(s, y, vis);
  double *w2 = (double *)malloc(sizeof(double) * 2 * (long)C * P * 2 * B * R);
#pragma omp parallel for collapse(2) schedule(dynamic)
  for (int c = 0; c < C; ++c)
    for (int p = 0; p < P; ++p) {
      const long cp = (long)c * P + p;
      counts[cp] = orc_rfi_null_cp(
          R, P, B, vis + (long)c * nbl * P * P * 2, p, w + cp * B * R * 2,
          detect ? -1 : fixed_k[cp], delta, shrink, w2 + cp * 2 * B * R * 2,
          evals + cp * R, evecs + cp * R * R * 2);
    }
  orc_beamform(s, y, w2, 2 * B, beams2b);
  (void)T;
  free(w2);
  return 0;
}
