/*
 * TEST INFRASTRUCTURE ONLY: C restatement of one Strang diffusion step of the
 * centred / M2 KRFFT pipeline, used as the CPU baseline that the GPU library
 * is timed against (never by the product).
 *
 *   q_out = E . Unfold IFFT3 Irecon( K . Recon FFT3 Fold(E . q_in) )
 *
 * Fold / unfold: src/centering_fold.jl:140-257, 481-580 (alive parity
 * channels of the stride-2 subgrid; all 8 for primitive groups, which is the
 * first radix-2 step of the M-grid FFT as in the M2 path).  Reconstruction:
 * F(h) = sum_a exp(-2 pi i h.t_a/N) G0(R_a^T h mod M), src/krfft.jl:529-545.
 * Inverse: G0(q) = (1/d) sum_alpha F(q + M alpha) through the spectral reverse
 * lookup, src/krfft.jl:1183-1311 (the weights, phase / d, come in the table).  The index/phase tables are built by
 * oracle/cstep.py from the NumPy oracle; this file only executes them.
 * FFT: Stockham radix 4/2/3 (lengths 2^a 3^b <= 64), FP64.  OpenMP over lines.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

typedef struct {
  int M[3], H[3], n_ch, offs[8][3], n_spec, d, nthreads;
  long H3, M3, n_alive;
  const int32_t* fidx;  /* n_spec x d: channel-linear index or -1 */
  const cplx* fph;      /* n_spec x d */
  const double* K;      /* n_spec */
  const int32_t* bidx;  /* n_alive x 8: spectral index or -1 */
  const cplx* bph;      /* n_alive x 8 (phase, applied after conjugation) */
  const uint8_t* bcj;   /* n_alive x 8 */
  double inv_d;         /* scale of the inverse reconstruction */
  cplx* ch;             /* n_alive */
  cplx* F;              /* n_spec */
  cplx* tw[3];          /* fold twiddles exp(-2 pi i n / M_d), n < M_d */
  cplx* ftw[3];         /* FFT twiddles exp(-2 pi i m / H_d) */
} cstep_plan;

/* Stockham autosort FFT of V interleaved lines (element j of line v at
 * x[j * V + v]), radices 4, 2, 3; tw[m] = exp(-2 pi i m / n).  Unnormalised;
 * sgn = -1 forward, +1 inverse.  Result in x; y is scratch of the same size. */
static void fft_v(int n, int V, int sgn, cplx* x, cplx* y, const cplx* tw) {
  cplx* src = x;
  cplx* dst = y;
  int Ns = 1;
  while (Ns < n) {
    const int rem = n / Ns;
    const int R = (rem % 4 == 0) ? 4 : (rem % 2 == 0 ? 2 : 3);
    const int nr = n / R;
    const int tstride = n / (Ns * R);
    for (int j = 0; j < nr; ++j) {
      const int jm = j % Ns;
      cplx w[4];
      for (int r = 0; r < R; ++r) {
        cplx t = tw[(jm * r * tstride) % n];
        w[r] = sgn > 0 ? conj(t) : t;
      }
      const int od = (j / Ns) * Ns * R + jm;
      for (int v = 0; v < V; ++v) {
        cplx a[4];
        for (int r = 0; r < R; ++r) a[r] = src[(j + r * nr) * V + v] * w[r];
        if (R == 2) {
          dst[od * V + v] = a[0] + a[1];
          dst[(od + Ns) * V + v] = a[0] - a[1];
        } else if (R == 4) {
          const cplx s02 = a[0] + a[2], d02 = a[0] - a[2], s13 = a[1] + a[3];
          const cplx d13 = (a[1] - a[3]) * (sgn < 0 ? -I : I);
          dst[od * V + v] = s02 + s13;
          dst[(od + Ns) * V + v] = d02 + d13;
          dst[(od + 2 * Ns) * V + v] = s02 - s13;
          dst[(od + 3 * Ns) * V + v] = d02 - d13;
        } else {
          const double s3 = 0.86602540378443864676 * (sgn < 0 ? -1.0 : 1.0);
          const cplx sa = a[1] + a[2], db = (a[1] - a[2]) * (I * s3);
          dst[od * V + v] = a[0] + sa;
          dst[(od + Ns) * V + v] = a[0] - 0.5 * sa + db;
          dst[(od + 2 * Ns) * V + v] = a[0] - 0.5 * sa - db;
        }
      }
    }
    cplx* t = src;
    src = dst;
    dst = t;
    Ns *= R;
  }
  if (src != x) memcpy(x, src, sizeof(cplx) * n * V);
}

enum { kVB = 4 };

static void fft3(cstep_plan* p, cplx* data, int sgn) {
  const int H0 = p->H[0], H1 = p->H[1], H2 = p->H[2];
  const long H3 = p->H3;
  /* axis 0: contiguous lines, one at a time */
  if (H0 > 1) {
    const long nl = (long)p->n_ch * H1 * H2;
#pragma omp parallel num_threads(p->nthreads)
    {
      cplx a[64], b[64];
#pragma omp for schedule(static)
      for (long l = 0; l < nl; ++l) {
        cplx* line = data + l * H0;
        memcpy(a, line, sizeof(cplx) * H0);
        fft_v(H0, 1, sgn, a, b, p->ftw[0]);
        memcpy(line, a, sizeof(cplx) * H0);
      }
    }
  }
  /* axes 1, 2: kVB neighbouring x-lines at a time */
  for (int ax = 1; ax < 3; ++ax) {
    const int n = p->H[ax];
    if (n == 1) continue;
    const int V = (H0 % kVB == 0) ? kVB : 1;
    const long s = ax == 1 ? H0 : (long)H0 * H1;
    const long nouter = ax == 1 ? (long)p->n_ch * H2 : (long)p->n_ch;
    const long ninner = ax == 1 ? H0 / V : (long)H0 * H1 / V;
#pragma omp parallel num_threads(p->nthreads)
    {
      cplx a[64 * kVB], b[64 * kVB];
#pragma omp for schedule(static)
      for (long l = 0; l < nouter * ninner; ++l) {
        const long o = l / ninner, i = l % ninner;
        cplx* base = ax == 1 ? data + (o / H2) * H3 + (o % H2) * (long)H0 * H1 + i * V
                             : data + o * H3 + i * V;
        for (int j = 0; j < n; ++j) memcpy(a + j * V, base + j * s, sizeof(cplx) * V);
        fft_v(n, V, sgn, a, b, p->ftw[ax]);
        for (int j = 0; j < n; ++j) memcpy(base + j * s, a + j * V, sizeof(cplx) * V);
      }
    }
  }
}

static inline long corner(const cstep_plan* p, int x, int y, int z, int e) {
  return (x + (e & 1) * p->H[0]) + (long)p->M[0] * ((y + ((e >> 1) & 1) * p->H[1]) +
                                                    (long)p->M[1] * (z + ((e >> 2) & 1) * p->H[2]));
}

static inline cplx fold_tw(const cstep_plan* p, int c, int x, int y, int z) {
  cplx t = 1.0;
  if (p->offs[c][0]) t *= p->tw[0][x];
  if (p->offs[c][1]) t *= p->tw[1][y];
  if (p->offs[c][2]) t *= p->tw[2][z];
  return t;
}

static void wht8(double* v) {
  for (int h = 1; h < 8; h <<= 1)
    for (int i = 0; i < 8; i += 2 * h)
      for (int j = i; j < i + h; ++j) {
        const double a = v[j], b = v[j + h];
        v[j] = a + b;
        v[j + h] = a - b;
      }
}

static int off_code(const cstep_plan* p, int c) { return p->offs[c][0] | (p->offs[c][1] << 1) | (p->offs[c][2] << 2); }

cstep_plan* cstep_create(const int32_t* M, int32_t n_ch, const int32_t* offs, int32_t n_spec, int32_t d,
                         const int32_t* fidx, const double* fph, const double* K, const int32_t* bidx,
                         const double* bph, const uint8_t* bcj, double bscale, int32_t nthreads) {
  cstep_plan* p = (cstep_plan*)calloc(1, sizeof(cstep_plan));
  for (int k = 0; k < 3; ++k) {
    p->M[k] = M[k];
    p->H[k] = M[k] / 2;
  }
  p->n_ch = n_ch;
  for (int c = 0; c < n_ch; ++c)
    for (int k = 0; k < 3; ++k) p->offs[c][k] = offs[3 * c + k];
  p->n_spec = n_spec;
  p->d = d;
  p->nthreads = nthreads > 0 ? nthreads : 1;
  p->H3 = (long)p->H[0] * p->H[1] * p->H[2];
  p->M3 = (long)M[0] * M[1] * M[2];
  p->n_alive = p->H3 * n_ch;
  p->fidx = fidx;
  p->fph = (const cplx*)fph;
  p->K = K;
  p->bidx = bidx;
  p->bph = (const cplx*)bph;
  p->bcj = bcj;
  p->inv_d = bscale;
  p->ch = (cplx*)malloc(sizeof(cplx) * p->n_alive);
  p->F = (cplx*)malloc(sizeof(cplx) * (n_spec > 0 ? n_spec : 1));
  for (int k = 0; k < 3; ++k) {
    p->tw[k] = (cplx*)malloc(sizeof(cplx) * M[k]);
    for (int n = 0; n < M[k]; ++n) p->tw[k][n] = cexp(-2.0 * M_PI * I * n / M[k]);
    p->ftw[k] = (cplx*)malloc(sizeof(cplx) * p->H[k]);
    for (int n = 0; n < p->H[k]; ++n) p->ftw[k][n] = cexp(-2.0 * M_PI * I * n / p->H[k]);
  }
  return p;
}

void cstep_destroy(cstep_plan* p) {
  if (!p) return;
  free(p->ch);
  free(p->F);
  for (int k = 0; k < 3; ++k) {
    free(p->tw[k]);
    free(p->ftw[k]);
  }
  free(p);
}

/* q_out = E D (E q_in); E may be NULL (pure diffusion). */
void cstep_step(cstep_plan* p, const double* q_in, double* q_out, const double* E) {
  const int H0 = p->H[0], H1 = p->H[1], H2 = p->H[2];
  const long H3 = p->H3;
  /* fold (+E) */
#pragma omp parallel for num_threads(p->nthreads) schedule(static)
  for (long k = 0; k < H3; ++k) {
    const int x = (int)(k % H0), y = (int)((k / H0) % H1), z = (int)(k / ((long)H0 * H1));
    double v[8];
    for (int e = 0; e < 8; ++e) {
      const long i = corner(p, x, y, z, e);
      v[e] = E ? q_in[i] * E[i] : q_in[i];
    }
    wht8(v);
    for (int c = 0; c < p->n_ch; ++c) p->ch[c * H3 + k] = v[off_code(p, c)] * fold_tw(p, c, x, y, z);
  }
  fft3(p, p->ch, -1);
  /* reconstruction + K */
#pragma omp parallel for num_threads(p->nthreads) schedule(static)
  for (long i = 0; i < p->n_spec; ++i) {
    cplx a = 0.0;
    for (int r = 0; r < p->d; ++r) {
      const int32_t li = p->fidx[i * p->d + r];
      if (li >= 0) a += p->fph[i * p->d + r] * p->ch[li];
    }
    p->F[i] = a * p->K[i];
  }
  /* inverse reconstruction */
#pragma omp parallel for num_threads(p->nthreads) schedule(static)
  for (long e = 0; e < p->n_alive; ++e) {
    cplx a = 0.0;
    for (int al = 0; al < 8; ++al) {
      const int32_t s = p->bidx[e * 8 + al];
      if (s < 0) continue;
      cplx f = p->F[s];
      if (p->bcj[e * 8 + al]) f = conj(f);
      a += f * p->bph[e * 8 + al];
    }
    p->ch[e] = a * p->inv_d;
  }
  fft3(p, p->ch, +1);
  /* unfold (+E): IFFT normalisation 1/H3 and inverse WHT 1/8 */
  const double sc = 1.0 / (8.0 * (double)H3);
#pragma omp parallel for num_threads(p->nthreads) schedule(static)
  for (long k = 0; k < H3; ++k) {
    const int x = (int)(k % H0), y = (int)((k / H0) % H1), z = (int)(k / ((long)H0 * H1));
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int c = 0; c < p->n_ch; ++c) v[off_code(p, c)] = creal(p->ch[c * H3 + k] * conj(fold_tw(p, c, x, y, z)));
    wht8(v);
    for (int e = 0; e < 8; ++e) {
      const long i = corner(p, x, y, z, e);
      const double r = v[e] * sc;
      q_out[i] = E ? r * E[i] : r;
    }
  }
}

void cstep_set_threads(cstep_plan* p, int32_t n) { p->nthreads = n > 0 ? n : 1; }

int cstep_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
