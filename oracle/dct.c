/*
 * TEST INFRASTRUCTURE -- part of the oracle.  Only tests/, the smoke check and the bench's CPU
 * baseline may call it; the product never does.
 *
 * 1-D cosine and sine transforms in the conventions of Ooura's fftsg package, which the
 * reference transcribes (ooura.rs:1-39, 158-255):
 *   dct2 (ddct, isgn = -1):  C[k] = sum_{j<n} a[j] cos(pi (j + 1/2) k / n)
 *   dct3 (ddct, isgn = +1):  C[k] = sum_{j<n} a[j] cos(pi j (k + 1/2) / n)
 *   dst3 (ddst, isgn = +1):  S[k] = sum_{j=1..n} A[j] sin(pi j (k + 1/2) / n), A[n] held in a[0]
 * and the three 2-D transforms the density solve calls (ooura.rs:71-98): rows of length n2
 * first, then columns of length n1, on a row-major n1 x n2 grid.
 *
 * The arithmetic is this file's own: Makhoul's DCT through a half-length complex FFT (radix-2,
 * twiddles from the C library in double, tabled once per length as Ooura's makewt/makect tables
 * are, ooura.rs:260-382), not Ooura's split-radix butterflies, so results agree
 * with the reference's to rounding, not bit for bit (the reference pins its own bits on macOS
 * only, ooura.rs:3124-3159).  Powers of two use the FFT, and so do 3 * 2^k (the half-length FFT
 * split by 3 in time, each third a power of two -- the reference itself takes powers of two only,
 * ooura.rs:16-17, so these sizes are pinned by the direct sums and scipy alone); any other length
 * falls back to the direct O(n^2) sum with double accumulation.
 *
 * The 2-D transforms take eight rows (then eight columns) at once, one in each lane of a GCC
 * vector, every lane through the one-sequence code's operations in the same order, so the bits
 * are those of the 1-D transforms applied line by line; the batched code is built twice, for
 * the baseline x86-64 and for AVX2, and runs as AVX2 where the host has it (no fused
 * multiply-add in either: -ffp-contract=off).  A 1024 x 1024 transform takes ~7 ms on one core,
 * as fast as a tabled library transform (scipy.fft's pocketfft, float32, one worker).
 *
 * Pinned by tests/test_oracle_dct.py: brute-force O(n^2) FP64 sums, scipy.fft, forward-inverse
 * identity, the single-cosine-mode eigenfunction of the Poisson solve, and the 2-D transforms
 * against the 1-D ones line by line, bit for bit, in both builds and with several threads.
 */
#include <alloca.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI 3.14159265358979323846

int orc_threads(void); /* oracle.c: the host threads the 2-D transforms may use */

static int is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }
/* the lengths the fast path transforms: 2^k, and 3 * 2^k (k >= 1: the half-length FFT then has a
 * radix-3 step) */
static int fast_len(int n) { return is_pow2(n) || (n % 6 == 0 && is_pow2(n / 6)); }

/* ---- generic body, instantiated for float (f32) and double (f64) ---------------------------- */
#define DEFINE_DCT(T, SFX)                                                                          \
    /* The transforms' twiddles of one length n (a power of two), computed once and kept: each     \
     * entry is the cosine or sine of the same double angle the transforms' expressions name       \
     * (2 pi k / len for the FFT's stage of length len, 2 pi k / n, pi k / 2n), rounded to T once.  \
     * fft: stage len = 2 h holds h pairs at [h - 1, 2 h - 1); wn: k <= n/2; w4n: k < n. */         \
    typedef struct {                                                                                \
        int n;                                                                                      \
        T *fft, *wn, *w4n;                                                                          \
    } tab_t_##SFX;                                                                                  \
    static tab_t_##SFX tabs_##SFX[64];                                                               \
    static void tab_fill_##SFX(tab_t_##SFX* t, int n);                                             \
    static const tab_t_##SFX* tab_##SFX(int n) {                                                   \
        int lg = 0;                                                                                 \
        while ((1 << lg) < n) ++lg;                                                                 \
        tab_t_##SFX* t = &tabs_##SFX[lg + (is_pow2(n) ? 0 : 32)]; /* 3 * 2^k beside the powers */ \
        /* (made once under a lock: the 2-D transforms' threads may ask for one at once; a made   \
         * table is seen through the acquire load of its length, stored last) */                  \
        if (__atomic_load_n(&t->n, __ATOMIC_ACQUIRE) == n) return t;                                \
        _Pragma("omp critical(orc_dct_tabs)")                                                       \
        if (t->n != n) tab_fill_##SFX(t, n);                                                        \
        return t;                                                                                   \
    }                                                                                               \
    static void tab_fill_##SFX(tab_t_##SFX* t, int n) {                                             \
        T* buf = (T*)malloc(sizeof(T) * (2 * (size_t)n + 2 * ((size_t)n / 2 + 1) + 2 * (size_t)n + 2)); \
        t->fft = buf;                                                                               \
        t->wn = buf + 2 * (size_t)n;                                                                \
        t->w4n = t->wn + 2 * ((size_t)n / 2 + 1);                                                   \
        for (int len = 2; len <= n; len <<= 1) {                                                    \
            int h = len >> 1;                                                                       \
            for (int k = 0; k < h; ++k) {                                                           \
                double ang = 2.0 * ORC_PI * k / len;                                                \
                t->fft[2 * (h - 1 + k)] = (T)cos(ang);                                              \
                t->fft[2 * (h - 1 + k) + 1] = (T)sin(ang);                                          \
            }                                                                                       \
        }                                                                                           \
        for (int k = 0; k <= n / 2; ++k) {                                                          \
            double ang = 2.0 * ORC_PI * k / n;                                                      \
            t->wn[2 * k] = (T)cos(ang);                                                             \
            t->wn[2 * k + 1] = (T)sin(ang);                                                         \
        }                                                                                           \
        for (int k = 0; k < n; ++k) {                                                               \
            double ang = ORC_PI * k / (2.0 * n);                                                    \
            t->w4n[2 * k] = (T)cos(ang);                                                            \
            t->w4n[2 * k + 1] = (T)sin(ang);                                                        \
        }                                                                                           \
        __atomic_store_n(&t->n, n, __ATOMIC_RELEASE);                                               \
    }                                                                                               \
    static void fft3_##SFX(int m, T* z, int sign);                                                \
    /* in-place radix-2 complex FFT of m points (re, im interleaved), sign -1 forward, +1 inverse  \
     * (no scaling); twiddles exp(sign 2 pi i k / len) from the length's table; m = 3 * 2^j goes   \
     * to fft3 */                                                                                   \
    static void fft_##SFX(int m, T* z, int sign) {                                                  \
        if (!is_pow2(m)) {                                                                          \
            fft3_##SFX(m, z, sign);                                                                 \
            return;                                                                                 \
        }                                                                                           \
        for (int i = 1, j = 0; i < m; ++i) {                                                        \
            int bit = m >> 1;                                                                       \
            for (; j & bit; bit >>= 1) j ^= bit;                                                    \
            j ^= bit;                                                                               \
            if (i < j) {                                                                            \
                T tr = z[2 * i], ti = z[2 * i + 1];                                                 \
                z[2 * i] = z[2 * j];                                                                \
                z[2 * i + 1] = z[2 * j + 1];                                                        \
                z[2 * j] = tr;                                                                      \
                z[2 * j + 1] = ti;                                                                  \
            }                                                                                       \
        }                                                                                           \
        const T* st = tab_##SFX(m)->fft;                                                            \
        /* the stage's twiddles with the sign on the sine, then its butterflies (each butterfly's  \
         * arithmetic is independent of the order they run in) */                                   \
        T* w = (T*)alloca(sizeof(T) * (size_t)m);                                                   \
        for (int len = 2; len <= m; len <<= 1) {                                                    \
            int h = len >> 1;                                                                       \
            for (int k = 0; k < h; ++k) {                                                           \
                w[2 * k] = st[2 * (h - 1 + k)];                                                     \
                w[2 * k + 1] = sign < 0 ? -st[2 * (h - 1 + k) + 1] : st[2 * (h - 1 + k) + 1];       \
            }                                                                                       \
            for (int s = 0; s < m; s += len) {                                                      \
                T* a = z + 2 * s;                                                                   \
                T* b = z + 2 * (s + h);                                                             \
                for (int k = 0; k < h; ++k) {                                                       \
                    const T wr = w[2 * k], wi = w[2 * k + 1];                                       \
                    T br = b[2 * k] * wr - b[2 * k + 1] * wi, bi = b[2 * k] * wi + b[2 * k + 1] * wr; \
                    b[2 * k] = a[2 * k] - br;                                                       \
                    b[2 * k + 1] = a[2 * k + 1] - bi;                                               \
                    a[2 * k] = a[2 * k] + br;                                                       \
                    a[2 * k + 1] = a[2 * k + 1] + bi;                                               \
                }                                                                                   \
            }                                                                                       \
        }                                                                                           \
    }                                                                                               \
    /* m = 3 p points, p a power of two: decimation in time by 3 -- the three p-point FFTs of      \
     * z[3j + r], then X[k] = Z0[k mod p] + w^k Z1[k mod p] + w^2k Z2[k mod p], w = e^{sign 2 pi i / m} \
     * (each twiddle the cosine and sine of the double angle, rounded to T) */                      \
    static void fft3_##SFX(int m, T* z, int sign) {                                                 \
        const int p = m / 3;                                                                        \
        T* s = (T*)malloc(sizeof(T) * 2 * (size_t)m);                                               \
        for (int r = 0; r < 3; ++r)                                                                 \
            for (int j = 0; j < p; ++j) {                                                           \
                s[2 * (r * p + j)] = z[2 * (3 * j + r)];                                            \
                s[2 * (r * p + j) + 1] = z[2 * (3 * j + r) + 1];                                    \
            }                                                                                       \
        for (int r = 0; r < 3; ++r) fft_##SFX(p, s + 2 * (size_t)r * p, sign);                      \
        for (int k = 0; k < m; ++k) {                                                               \
            const int kk = k % p;                                                                   \
            T xr = s[2 * kk], xi = s[2 * kk + 1];                                                   \
            for (int r = 1; r < 3; ++r) {                                                           \
                const double ang = 2.0 * ORC_PI * (double)((long)r * k % m) / m;                    \
                const T wr = (T)cos(ang), wi = (T)(sign * sin(ang));                                \
                const T br = s[2 * (r * p + kk)], bi = s[2 * (r * p + kk) + 1];                     \
                xr += wr * br - wi * bi;                                                            \
                xi += wr * bi + wi * br;                                                            \
            }                                                                                       \
            z[2 * k] = xr;                                                                          \
            z[2 * k + 1] = xi;                                                                      \
        }                                                                                           \
        free(s);                                                                                    \
    }                                                                                               \
    /* DFT V[k] = sum_m v[m] e^{-2 pi i m k / n}, k = 0..n/2, of a real sequence v (n even),        \
     * through an n/2-point complex FFT; out holds n/2 + 1 complex values */                        \
    static void rfft_##SFX(int n, const T* v, T* out, T* z) {                                       \
        int h = n / 2;                                                                              \
        for (int m = 0; m < h; ++m) {                                                               \
            z[2 * m] = v[2 * m];                                                                    \
            z[2 * m + 1] = v[2 * m + 1];                                                            \
        }                                                                                           \
        fft_##SFX(h, z, -1);                                                                        \
        for (int k = 0; k <= h; ++k) {                                                              \
            int k1 = k % h, k2 = (h - k) % h;                                                       \
            T ar = z[2 * k1], ai = z[2 * k1 + 1], br = z[2 * k2], bi = -z[2 * k2 + 1];              \
            T er = (ar + br) * (T)0.5, ei = (ai + bi) * (T)0.5; /* even part */                     \
            T orr = (ar - br) * (T)0.5, oi = (ai - bi) * (T)0.5; /* i * odd part */                 \
            const T* wn = tab_##SFX(n)->wn; /* cos, sin of 2 pi k / n */                            \
            T wr = wn[2 * k], wi = -wn[2 * k + 1];                                                  \
            /* V = E + w * O, O = -i * (orr + i oi) = oi - i orr */                                 \
            T xr = oi, xi = -orr;                                                                   \
            out[2 * k] = er + (wr * xr - wi * xi);                                                  \
            out[2 * k + 1] = ei + (wr * xi + wi * xr);                                              \
        }                                                                                           \
    }                                                                                               \
    static void direct_dct2_##SFX(int n, T* a) {                                                    \
        double* t = (double*)malloc(sizeof(double) * n);                                            \
        for (int k = 0; k < n; ++k) {                                                               \
            double s = 0.0;                                                                         \
            for (int j = 0; j < n; ++j) s += (double)a[j] * cos(ORC_PI * (j + 0.5) * k / n);        \
            t[k] = s;                                                                               \
        }                                                                                           \
        for (int k = 0; k < n; ++k) a[k] = (T)t[k];                                                 \
        free(t);                                                                                    \
    }                                                                                               \
    static void direct_dct3_##SFX(int n, T* a) {                                                    \
        double* t = (double*)malloc(sizeof(double) * n);                                            \
        for (int k = 0; k < n; ++k) {                                                               \
            double s = 0.0;                                                                         \
            for (int j = 0; j < n; ++j) s += (double)a[j] * cos(ORC_PI * j * (k + 0.5) / n);        \
            t[k] = s;                                                                               \
        }                                                                                           \
        for (int k = 0; k < n; ++k) a[k] = (T)t[k];                                                 \
        free(t);                                                                                    \
    }                                                                                               \
    /* Ooura ddct isgn = -1 */                                                                      \
    void orc_dct2_##SFX(int n, T* a) {                                                              \
        if (n == 1) return;                                                                         \
        if (!fast_len(n)) {                                                                          \
            direct_dct2_##SFX(n, a);                                                                \
            return;                                                                                 \
        }                                                                                           \
        T* v = (T*)malloc(sizeof(T) * (5 * n + 4));                                                 \
        T* V = v + n;                                                                               \
        T* z = V + n + 2;                                                                           \
        for (int k = 0; k < n / 2; ++k) {                                                           \
            v[k] = a[2 * k];                                                                        \
            v[n - 1 - k] = a[2 * k + 1];                                                            \
        }                                                                                           \
        rfft_##SFX(n, v, V, z);                                                                     \
        for (int k = 0; k < n; ++k) {                                                               \
            T vr, vi;                                                                               \
            if (k <= n / 2) {                                                                       \
                vr = V[2 * k];                                                                      \
                vi = V[2 * k + 1];                                                                  \
            } else {                                                                                \
                vr = V[2 * (n - k)];                                                                \
                vi = -V[2 * (n - k) + 1];                                                           \
            }                                                                                       \
            const T* w4 = tab_##SFX(n)->w4n; /* cos, sin of pi k / 2n */                            \
            a[k] = w4[2 * k] * vr - (-w4[2 * k + 1]) * vi;                                          \
        }                                                                                           \
        free(v);                                                                                    \
    }                                                                                               \
    /* Ooura ddct isgn = +1: y[m] = sum_k a[k] cos(pi k (2m+1) / 2n).  With V[k] =                  \
     * e^{i pi k / 2n} (c[k] - i c[n-k]), c = a except c[0] = 2 a[0], c[n] = 0, the real sequence   \
     * v = IDFT(V) / 2 gives y[2m] = v[m], y[2m+1] = v[n-1-m]. */                                  \
    void orc_dct3_##SFX(int n, T* a) {                                                              \
        if (n == 1) return;                                                                         \
        if (!fast_len(n)) {                                                                          \
            direct_dct3_##SFX(n, a);                                                                \
            return;                                                                                 \
        }                                                                                           \
        int h = n / 2;                                                                              \
        T* buf = (T*)malloc(sizeof(T) * (4 * n + 8));                                               \
        T* V = buf;           /* n/2 + 1 complex */                                                 \
        T* z = V + n + 2;     /* n/2 complex */                                                     \
        for (int k = 0; k <= h; ++k) {                                                              \
            T ck = (k == 0) ? (T)2 * a[0] : a[k % n];                                               \
            T cnk = (k == 0) ? (T)0 : a[n - k];                                                     \
            if (k == h) ck = a[h], cnk = a[h];                                                      \
            const T* w4 = tab_##SFX(n)->w4n;                                                        \
            T wr = w4[2 * k], wi = w4[2 * k + 1];                                                   \
            V[2 * k] = wr * ck + wi * cnk;                                                          \
            V[2 * k + 1] = wi * ck - wr * cnk;                                                      \
        }                                                                                           \
        /* inverse real DFT via an n/2-point complex IFFT:                                          \
         * Z[k] = (V[k] + conj(V[h-k])) + i e^{2 pi i k/n} (V[k] - conj(V[h-k])) */                 \
        for (int k = 0; k < h; ++k) {                                                               \
            T ar = V[2 * k], ai = V[2 * k + 1], br = V[2 * (h - k)], bi = -V[2 * (h - k) + 1];      \
            T er = ar + br, ei = ai + bi, dr = ar - br, di = ai - bi;                               \
            const T* wn = tab_##SFX(n)->wn;                                                         \
            T wr = wn[2 * k], wi = wn[2 * k + 1];                                                   \
            T tr = wr * dr - wi * di, ti = wr * di + wi * dr; /* w * d */                           \
            z[2 * k] = er - ti;                                                                     \
            z[2 * k + 1] = ei + tr;                                                                 \
        }                                                                                           \
        fft_##SFX(h, z, +1);                                                                        \
        /* v[2m] = Re z[m] / n, v[2m+1] = Im z[m] / n; y = v * n / 2 */                             \
        T* y = (T*)malloc(sizeof(T) * n);                                                           \
        for (int m = 0; m < h; ++m) {                                                               \
            T v0 = z[2 * m] * (T)0.5, v1 = z[2 * m + 1] * (T)0.5;                                   \
            y[2 * m] = v0;                                                                          \
            y[2 * m + 1] = v1;                                                                      \
        }                                                                                           \
        for (int m = 0; m < h; ++m) {                                                               \
            a[2 * m] = y[m];                                                                        \
            a[2 * m + 1] = y[n - 1 - m];                                                            \
        }                                                                                           \
        free(y);                                                                                    \
        free(buf);                                                                                  \
    }                                                                                               \
    /* Ooura ddst isgn = +1: S[k] = (-1)^k dct3(b)[k], b[0] = a[0] (= A[n]), b[j] = a[n-j] */       \
    void orc_dst3_##SFX(int n, T* a) {                                                              \
        if (n == 1) return;                                                                         \
        for (int j = 1; j < n - j; ++j) {                                                           \
            T t = a[j];                                                                             \
            a[j] = a[n - j];                                                                        \
            a[n - j] = t;                                                                           \
        }                                                                                           \
        orc_dct3_##SFX(n, a);                                                                       \
        for (int k = 1; k < n; k += 2) a[k] = -a[k];                                                \
    }                                                                                               \
    /* ---- batched transforms: BL sequences at once, one in each lane of a vector.  Every lane goes \
     * through the operations of the one-sequence code above in the same order (the same rounded    \
     * operations on the same constants), so each lane's result is that code's bit for bit; the     \
     * 2-D transforms use them, rows and columns BL at a time (tests/test_oracle_dct.py compares    \
     * them with the one-sequence transforms row by row and column by column). */                   \
    typedef T bv_##SFX __attribute__((vector_size(8 * sizeof(T))));                                 \
    enum { BL_##SFX = 8 };                                                                          \
    /* one host thread's scratch: the sequences (len), the transforms' buffer (4 len + 2), the      \
     * twiddles (len scalars) */                                                                    \
    static bv_##SFX* bscratch_##SFX(int len, T** w) {                                               \
        size_t bytes = (5 * (size_t)len + 2) * sizeof(bv_##SFX) + (size_t)len * sizeof(T);          \
        bytes = (bytes + 63) / 64 * 64;                                                             \
        bv_##SFX* q = (bv_##SFX*)aligned_alloc(64, bytes);                                          \
        *w = (T*)(q + 5 * (size_t)len + 2);                                                         \
        return q;                                                                                   \
    }                                                                                               \
    /* a length the batched code does not take (no fast transform) goes one sequence at a time */   \
    enum { B_##SFX = 16 };                                                                          \
    static void colblock_##SFX(int n1, int n2, T* a, T* t, int j0, void (*col)(int, T*)) {          \
        const int nb = n2 - j0 < B_##SFX ? n2 - j0 : B_##SFX;                                       \
        for (int i = 0; i < n1; ++i)                                                                \
            for (int q = 0; q < nb; ++q) t[(size_t)q * n1 + i] = a[(size_t)i * n2 + j0 + q];        \
        for (int q = 0; q < nb; ++q) col(n1, t + (size_t)q * n1);                                   \
        for (int i = 0; i < n1; ++i)                                                                \
            for (int q = 0; q < nb; ++q) a[(size_t)i * n2 + j0 + q] = t[(size_t)q * n1 + i];        \
    }

/* The batched transforms of one element type, built once per instruction set (TAG). */
#define DEFINE_BATCH(T, SFX, TAG)                                                                   \
    static void bfft3_##SFX##_##TAG(int m, bv_##SFX* z, int sign, bv_##SFX* s, T* w);               \
    /* fft on BL sequences (w: m scalars of twiddle scratch, s3: 2 m vectors for m = 3 * 2^j) */    \
    static void bfft_##SFX##_##TAG(int m, bv_##SFX* z, int sign, bv_##SFX* s3, T* w) {              \
        if (!is_pow2(m)) {                                                                          \
            bfft3_##SFX##_##TAG(m, z, sign, s3, w);                                                 \
            return;                                                                                 \
        }                                                                                           \
        for (int i = 1, j = 0; i < m; ++i) {                                                        \
            int bit = m >> 1;                                                                       \
            for (; j & bit; bit >>= 1) j ^= bit;                                                    \
            j ^= bit;                                                                               \
            if (i < j) {                                                                            \
                bv_##SFX tr = z[2 * i], ti = z[2 * i + 1];                                          \
                z[2 * i] = z[2 * j];                                                                \
                z[2 * i + 1] = z[2 * j + 1];                                                        \
                z[2 * j] = tr;                                                                      \
                z[2 * j + 1] = ti;                                                                  \
            }                                                                                       \
        }                                                                                           \
        const T* st = tab_##SFX(m)->fft;                                                            \
        for (int len = 2; len <= m; len <<= 1) {                                                    \
            int h = len >> 1;                                                                       \
            for (int k = 0; k < h; ++k) {                                                           \
                w[2 * k] = st[2 * (h - 1 + k)];                                                     \
                w[2 * k + 1] = sign < 0 ? -st[2 * (h - 1 + k) + 1] : st[2 * (h - 1 + k) + 1];       \
            }                                                                                       \
            for (int s = 0; s < m; s += len) {                                                      \
                bv_##SFX* a = z + 2 * s;                                                            \
                bv_##SFX* b = z + 2 * (s + h);                                                      \
                for (int k = 0; k < h; ++k) {                                                       \
                    const T wr = w[2 * k], wi = w[2 * k + 1];                                       \
                    bv_##SFX br = b[2 * k] * wr - b[2 * k + 1] * wi, bi = b[2 * k] * wi + b[2 * k + 1] * wr; \
                    b[2 * k] = a[2 * k] - br;                                                       \
                    b[2 * k + 1] = a[2 * k + 1] - bi;                                               \
                    a[2 * k] = a[2 * k] + br;                                                       \
                    a[2 * k + 1] = a[2 * k + 1] + bi;                                               \
                }                                                                                   \
            }                                                                                       \
        }                                                                                           \
    }                                                                                               \
    static void bfft3_##SFX##_##TAG(int m, bv_##SFX* z, int sign, bv_##SFX* s, T* w) {              \
        const int p = m / 3;                                                                        \
        for (int r = 0; r < 3; ++r)                                                                 \
            for (int j = 0; j < p; ++j) {                                                           \
                s[2 * (r * p + j)] = z[2 * (3 * j + r)];                                            \
                s[2 * (r * p + j) + 1] = z[2 * (3 * j + r) + 1];                                    \
            }                                                                                       \
        for (int r = 0; r < 3; ++r) bfft_##SFX##_##TAG(p, s + 2 * (size_t)r * p, sign, NULL, w);    \
        for (int k = 0; k < m; ++k) {                                                               \
            const int kk = k % p;                                                                   \
            bv_##SFX xr = s[2 * kk], xi = s[2 * kk + 1];                                            \
            for (int r = 1; r < 3; ++r) {                                                           \
                const double ang = 2.0 * ORC_PI * (double)((long)r * k % m) / m;                    \
                const T wr = (T)cos(ang), wi = (T)(sign * sin(ang));                                \
                const bv_##SFX br = s[2 * (r * p + kk)], bi = s[2 * (r * p + kk) + 1];              \
                xr += wr * br - wi * bi;                                                            \
                xi += wr * bi + wi * br;                                                            \
            }                                                                                       \
            z[2 * k] = xr;                                                                          \
            z[2 * k + 1] = xi;                                                                      \
        }                                                                                           \
    }                                                                                               \
    static void brfft_##SFX##_##TAG(int n, const bv_##SFX* v, bv_##SFX* out, bv_##SFX* z, bv_##SFX* s3, T* w) { \
        const int h = n / 2;                                                                        \
        for (int m = 0; m < n; ++m) z[m] = v[m];                                                    \
        bfft_##SFX##_##TAG(h, z, -1, s3, w);                                                        \
        const T* wn = tab_##SFX(n)->wn;                                                             \
        for (int k = 0; k <= h; ++k) {                                                              \
            int k1 = k % h, k2 = (h - k) % h;                                                       \
            bv_##SFX ar = z[2 * k1], ai = z[2 * k1 + 1], br = z[2 * k2], bi = -z[2 * k2 + 1];       \
            bv_##SFX er = (ar + br) * (T)0.5, ei = (ai + bi) * (T)0.5;                              \
            bv_##SFX orr = (ar - br) * (T)0.5, oi = (ai - bi) * (T)0.5;                             \
            T wr = wn[2 * k], wi = -wn[2 * k + 1];                                                  \
            bv_##SFX xr = oi, xi = -orr;                                                            \
            out[2 * k] = er + (wr * xr - wi * xi);                                                  \
            out[2 * k + 1] = ei + (wr * xi + wi * xr);                                              \
        }                                                                                           \
    }                                                                                               \
    /* buf: 4 n + 2 vectors */                                                                      \
    static void bdct2_##SFX##_##TAG(int n, bv_##SFX* a, bv_##SFX* buf, T* w) {                      \
        bv_##SFX* v = buf;                                                                          \
        bv_##SFX* V = v + n;                                                                        \
        bv_##SFX* z = V + n + 2;                                                                    \
        bv_##SFX* s3 = z + n;                                                                       \
        for (int k = 0; k < n / 2; ++k) {                                                           \
            v[k] = a[2 * k];                                                                        \
            v[n - 1 - k] = a[2 * k + 1];                                                            \
        }                                                                                           \
        brfft_##SFX##_##TAG(n, v, V, z, s3, w);                                                     \
        const T* w4 = tab_##SFX(n)->w4n;                                                            \
        for (int k = 0; k < n; ++k) {                                                               \
            bv_##SFX vr, vi;                                                                        \
            if (k <= n / 2) {                                                                       \
                vr = V[2 * k];                                                                      \
                vi = V[2 * k + 1];                                                                  \
            } else {                                                                                \
                vr = V[2 * (n - k)];                                                                \
                vi = -V[2 * (n - k) + 1];                                                           \
            }                                                                                       \
            a[k] = w4[2 * k] * vr - (-w4[2 * k + 1]) * vi;                                          \
        }                                                                                           \
    }                                                                                               \
    static void bdct3_##SFX##_##TAG(int n, bv_##SFX* a, bv_##SFX* buf, T* w) {                      \
        const int h = n / 2;                                                                        \
        bv_##SFX* V = buf;                                                                          \
        bv_##SFX* z = V + n + 2;                                                                    \
        bv_##SFX* y = z + n;                                                                        \
        bv_##SFX* s3 = y + n;                                                                       \
        const bv_##SFX zero = {0};                                                                  \
        const T* w4 = tab_##SFX(n)->w4n;                                                            \
        for (int k = 0; k <= h; ++k) {                                                              \
            bv_##SFX ck = (k == 0) ? (T)2 * a[0] : a[k % n];                                        \
            bv_##SFX cnk = (k == 0) ? zero : a[n - k];                                              \
            if (k == h) ck = a[h], cnk = a[h];                                                      \
            const T wr = w4[2 * k], wi = w4[2 * k + 1];                                             \
            V[2 * k] = wr * ck + wi * cnk;                                                          \
            V[2 * k + 1] = wi * ck - wr * cnk;                                                      \
        }                                                                                           \
        const T* wn = tab_##SFX(n)->wn;                                                             \
        for (int k = 0; k < h; ++k) {                                                               \
            bv_##SFX ar = V[2 * k], ai = V[2 * k + 1], br = V[2 * (h - k)], bi = -V[2 * (h - k) + 1]; \
            bv_##SFX er = ar + br, ei = ai + bi, dr = ar - br, di = ai - bi;                        \
            const T wr = wn[2 * k], wi = wn[2 * k + 1];                                             \
            bv_##SFX tr = wr * dr - wi * di, ti = wr * di + wi * dr;                                \
            z[2 * k] = er - ti;                                                                     \
            z[2 * k + 1] = ei + tr;                                                                 \
        }                                                                                           \
        bfft_##SFX##_##TAG(h, z, +1, s3, w);                                                        \
        for (int m = 0; m < h; ++m) {                                                               \
            y[2 * m] = z[2 * m] * (T)0.5;                                                           \
            y[2 * m + 1] = z[2 * m + 1] * (T)0.5;                                                   \
        }                                                                                           \
        for (int m = 0; m < h; ++m) {                                                               \
            a[2 * m] = y[m];                                                                        \
            a[2 * m + 1] = y[n - 1 - m];                                                            \
        }                                                                                           \
    }                                                                                               \
    static void bdst3_##SFX##_##TAG(int n, bv_##SFX* a, bv_##SFX* buf, T* w) {                      \
        for (int j = 1; j < n - j; ++j) {                                                           \
            bv_##SFX t = a[j];                                                                      \
            a[j] = a[n - j];                                                                        \
            a[n - j] = t;                                                                           \
        }                                                                                           \
        bdct3_##SFX##_##TAG(n, a, buf, w);                                                          \
        for (int k = 1; k < n; k += 2) a[k] = -a[k];                                                \
    }                                                                                               \
    /* kind 0 dct2, 1 dct3, 2 dst3 of BL sequences of length n (n a fast length) */                 \
    static void bxf_##SFX##_##TAG(int kind, int n, bv_##SFX* a, bv_##SFX* buf, T* w) {              \
        if (n == 1) return;                                                                         \
        if (kind == 0)                                                                              \
            bdct2_##SFX##_##TAG(n, a, buf, w);                                                      \
        else if (kind == 1)                                                                         \
            bdct3_##SFX##_##TAG(n, a, buf, w);                                                      \
        else                                                                                        \
            bdst3_##SFX##_##TAG(n, a, buf, w);                                                      \
    }                                                                                               \
    /* rows i0 .. i0 + BL - 1 (those below n1) in the lanes of q, transformed, and back */          \
    static void brows_##SFX##_##TAG(int n1, int n2, T* a, int i0, int kind, bv_##SFX* q, bv_##SFX* buf, T* w) { \
        const int nb = n1 - i0 < BL_##SFX ? n1 - i0 : BL_##SFX;                                     \
        for (int j = 0; j < n2; ++j) {                                                              \
            bv_##SFX x = {0};                                                                       \
            for (int l = 0; l < nb; ++l) x[l] = a[(size_t)(i0 + l) * n2 + j];                       \
            q[j] = x;                                                                               \
        }                                                                                           \
        bxf_##SFX##_##TAG(kind, n2, q, buf, w);                                                     \
        for (int l = 0; l < nb; ++l)                                                                \
            for (int j = 0; j < n2; ++j) a[(size_t)(i0 + l) * n2 + j] = q[j][l];                    \
    }                                                                                               \
    /* columns j0 .. j0 + BL - 1 (those below n2): each row's run of BL values is one vector */     \
    static void bcols_##SFX##_##TAG(int n1, int n2, T* a, int j0, int kind, bv_##SFX* q, bv_##SFX* buf, T* w) { \
        const int nb = n2 - j0 < BL_##SFX ? n2 - j0 : BL_##SFX;                                     \
        for (int i = 0; i < n1; ++i) {                                                              \
            bv_##SFX x = {0};                                                                       \
            if (nb == BL_##SFX)                                                                     \
                memcpy(&x, a + (size_t)i * n2 + j0, sizeof x);                                      \
            else                                                                                    \
                for (int l = 0; l < nb; ++l) x[l] = a[(size_t)i * n2 + j0 + l];                     \
            q[i] = x;                                                                               \
        }                                                                                           \
        bxf_##SFX##_##TAG(kind, n1, q, buf, w);                                                     \
        for (int i = 0; i < n1; ++i) {                                                              \
            if (nb == BL_##SFX)                                                                     \
                memcpy(a + (size_t)i * n2 + j0, &q[i], sizeof q[i]);                                \
            else                                                                                    \
                for (int l = 0; l < nb; ++l) a[(size_t)i * n2 + j0 + l] = q[i][l];                  \
        }                                                                                           \
    }

#define DEFINE_2D(T, SFX)                                                                           \
    /* 2-D: rows (length n2) with `row` (kind rk), then columns (length n1) with `col` (kind ck);   \
     * groups of rows, and of columns, are independent: with several host threads each takes its own. \
     * The batched code runs as built for AVX2 where the host has it, else for the baseline x86-64. */ \
    static void xt2d_##SFX(int n1, int n2, T* a, void (*row)(int, T*), void (*col)(int, T*), int rk, int ck) { \
        const int nt = orc_threads();                                                               \
        const int avx2 = batch_isa();                                                               \
        const int len = n1 > n2 ? n1 : n2;                                                          \
        const int fr = fast_len(n2), fc = fast_len(n1);                                             \
        _Pragma("omp parallel num_threads(nt) if (nt > 1)")                                         \
        {                                                                                           \
            T* w;                                                                                   \
            bv_##SFX* q = bscratch_##SFX(len, &w);                                                  \
            bv_##SFX* buf = q + len;                                                                \
            if (fr) {                                                                               \
                _Pragma("omp for schedule(static)")                                                 \
                for (int i0 = 0; i0 < n1; i0 += BL_##SFX) {                                         \
                    if (avx2)                                                                       \
                        brows_##SFX##_avx2(n1, n2, a, i0, rk, q, buf, w);                           \
                    else                                                                            \
                        brows_##SFX##_base(n1, n2, a, i0, rk, q, buf, w);                           \
                }                                                                                   \
            } else {                                                                                \
                _Pragma("omp for schedule(static)")                                                 \
                for (int i = 0; i < n1; ++i) row(n2, a + (size_t)i * n2);                           \
            }                                                                                       \
            if (fc) {                                                                               \
                _Pragma("omp for schedule(static)")                                                 \
                for (int j0 = 0; j0 < n2; j0 += BL_##SFX) {                                         \
                    if (avx2)                                                                       \
                        bcols_##SFX##_avx2(n1, n2, a, j0, ck, q, buf, w);                           \
                    else                                                                            \
                        bcols_##SFX##_base(n1, n2, a, j0, ck, q, buf, w);                           \
                }                                                                                   \
            } else {                                                                                \
                T* t = (T*)malloc(sizeof(T) * (size_t)n1 * B_##SFX);                                \
                _Pragma("omp for schedule(static)")                                                 \
                for (int j0 = 0; j0 < n2; j0 += B_##SFX) colblock_##SFX(n1, n2, a, t, j0, col);     \
                free(t);                                                                            \
            }                                                                                       \
            free(q);                                                                                \
        }                                                                                           \
    }                                                                                               \
    void orc_ddct2d_fwd_##SFX(int n1, int n2, T* a) { xt2d_##SFX(n1, n2, a, orc_dct2_##SFX, orc_dct2_##SFX, 0, 0); } \
    void orc_ddct2d_inv_##SFX(int n1, int n2, T* a) { xt2d_##SFX(n1, n2, a, orc_dct3_##SFX, orc_dct3_##SFX, 1, 1); } \
    /* ddsct2d(+1): cosine along the rows, sine along the columns */                                \
    void orc_ddsct2d_inv_##SFX(int n1, int n2, T* a) { xt2d_##SFX(n1, n2, a, orc_dct3_##SFX, orc_dst3_##SFX, 1, 2); } \
    /* ddcst2d(+1): sine along the rows, cosine along the columns */                                \
    void orc_ddcst2d_inv_##SFX(int n1, int n2, T* a) { xt2d_##SFX(n1, n2, a, orc_dst3_##SFX, orc_dct3_##SFX, 2, 1); }

DEFINE_DCT(float, f32)
DEFINE_DCT(double, f64)

/* -1: the batched code for the host's CPU (AVX2 where it has it), 0: the baseline x86-64 build */
static int g_batch_isa = -1;
void orc_set_batch_isa(int isa) { g_batch_isa = isa; }
static int batch_isa(void) { return g_batch_isa != 0 && __builtin_cpu_supports("avx2"); }

DEFINE_BATCH(float, f32, base)
DEFINE_BATCH(double, f64, base)
#pragma GCC push_options
#pragma GCC target("avx2")
DEFINE_BATCH(float, f32, avx2)
DEFINE_BATCH(double, f64, avx2)
#pragma GCC pop_options
DEFINE_2D(float, f32)
DEFINE_2D(double, f64)

/* Direct O(n^2) FP64 references, for pinning the fast transforms. */
void orc_direct_dct2_f64(int n, double* a) { direct_dct2_f64(n, a); }
void orc_direct_dct3_f64(int n, double* a) { direct_dct3_f64(n, a); }
void orc_direct_dst3_f64(int n, double* a) {
    double* t = (double*)malloc(sizeof(double) * n);
    for (int k = 0; k < n; ++k) {
        double s = 0.0;
        for (int j = 1; j <= n; ++j) {
            double A = (j == n) ? a[0] : a[j];
            s += A * sin(ORC_PI * j * (k + 0.5) / n);
        }
        t[k] = s;
    }
    memcpy(a, t, sizeof(double) * n);
    free(t);
}
