/*
 * orc_impl.h — CPU oracle of Hammerhead.jl's WIDIM correlation path
 * (TEST INFRASTRUCTURE; see orc_common.h). Included twice by oracle.c:
 *   REAL = double, SFX = _f64  : the FP64 oracle (the reference run with
 *                                Float64 images);
 *   REAL = float,  SFX = _f32  : the reference's Float32 pipeline, keeping
 *                                its Float64 islands (moment sums, UQ sums,
 *                                substitution/replacement medians, the UOD
 *                                residual division).
 * Every function cites the reference file:line it restates (paths relative
 * to the Hammerhead.jl source tree). Arrays are column-major, 0-based in
 * memory; positions exchanged with callers are 1-based like the reference.
 */

#define F(name) ORC_CAT(name, SFX)
#define RLOG(x) ((REAL)log((double)(x)))
#define REAL_EPS ((REAL)(sizeof(REAL) == 4 ? 1.1920928955078125e-07 : 2.220446049250313e-16))

/* ======================================================================
 * FFT: unnormalized forward (e^{-2 pi i k n/N}) and backward transforms, as
 * FFTW's plan_fft!/plan_bfft! (src/correlators.jl:68-71). Radix-2 for
 * powers of two, direct DFT otherwise. Twiddles from double cos/sin.
 * ==================================================================== */
typedef struct {
    int n;
    int pow2;
    REAL *wr, *wi; /* e^{-2 pi i k/n}, k < n */
} F(orc_fftplan);

static void F(orc_fftplan_init)(F(orc_fftplan) * p, int n)
{
    p->n = n;
    p->pow2 = n > 0 && (n & (n - 1)) == 0;
    p->wr = (REAL *)malloc(sizeof(REAL) * n);
    p->wi = (REAL *)malloc(sizeof(REAL) * n);
    for (int k = 0; k < n; k++) {
        double a = -2.0 * M_PI * (double)k / (double)n;
        p->wr[k] = (REAL)cos(a);
        p->wi[k] = (REAL)sin(a);
    }
}

static void F(orc_fftplan_free)(F(orc_fftplan) * p)
{
    free(p->wr);
    free(p->wi);
}

/* in-place 1D transform of a contiguous line; inv = 1 for e^{+...} */
static void F(orc_fft_line)(const F(orc_fftplan) * p, REAL *re, REAL *im, int inv, REAL *tr, REAL *ti)
{
    int n = p->n;
    if (p->pow2) {
        for (int i = 1, j = 0; i < n; i++) {
            int bit = n >> 1;
            for (; j & bit; bit >>= 1) j ^= bit;
            j ^= bit;
            if (i < j) {
                REAL t = re[i]; re[i] = re[j]; re[j] = t;
                t = im[i]; im[i] = im[j]; im[j] = t;
            }
        }
        for (int len = 2; len <= n; len <<= 1) {
            int half = len >> 1, step = n / len;
            for (int i = 0; i < n; i += len) {
                for (int k = 0; k < half; k++) {
                    REAL w_r = p->wr[k * step];
                    REAL w_i = inv ? -p->wi[k * step] : p->wi[k * step];
                    int a = i + k, b = i + k + half;
                    REAL xr = re[b] * w_r - im[b] * w_i;
                    REAL xi = re[b] * w_i + im[b] * w_r;
                    re[b] = re[a] - xr;
                    im[b] = im[a] - xi;
                    re[a] = re[a] + xr;
                    im[a] = im[a] + xi;
                }
            }
        }
    } else {
        for (int k = 0; k < n; k++) {
            REAL sr = 0, si = 0;
            for (int j = 0; j < n; j++) {
                int idx = (int)(((long)k * j) % n);
                REAL w_r = p->wr[idx];
                REAL w_i = inv ? -p->wi[idx] : p->wi[idx];
                sr += re[j] * w_r - im[j] * w_i;
                si += re[j] * w_i + im[j] * w_r;
            }
            tr[k] = sr;
            ti[k] = si;
        }
        memcpy(re, tr, sizeof(REAL) * n);
        memcpy(im, ti, sizeof(REAL) * n);
    }
}

/* 2D transform over dims (1, 2) of an n1 x n2 column-major complex array. */
static void F(orc_fft2)(const F(orc_fftplan) * p1, const F(orc_fftplan) * p2, REAL *re, REAL *im,
                        int inv, REAL *lr, REAL *li, REAL *tr, REAL *ti)
{
    int n1 = p1->n, n2 = p2->n;
    for (int j = 0; j < n2; j++) F(orc_fft_line)(p1, re + (size_t)j * n1, im + (size_t)j * n1, inv, tr, ti);
    for (int i = 0; i < n1; i++) {
        for (int j = 0; j < n2; j++) {
            lr[j] = re[i + (size_t)j * n1];
            li[j] = im[i + (size_t)j * n1];
        }
        F(orc_fft_line)(p2, lr, li, inv, tr, ti);
        for (int j = 0; j < n2; j++) {
            re[i + (size_t)j * n1] = lr[j];
            im[i + (size_t)j * n1] = li[j];
        }
    }
}

/* ======================================================================
 * Correlator (src/correlators.jl:42-181)
 * ==================================================================== */
typedef struct {
    int wr, wc, m1, m2;
    int sr, sc; /* search area (frame B): the window size unless enlarged */
    int method, padding;
    REAL *apod; /* wr x wc                                  */
    REAL *gain; /* m1 x m2, NULL when not padded            */
    REAL *W;    /* phase filter m1 x m2 (unshifted), or NULL */
    F(orc_fftplan) p1, p2;
} F(orc_corr);

/* apodization_window (src/correlators.jl:18-25): computed in Float64,
 * stored as T. */
void F(orc_apodization)(int wr, int wc, int apod, REAL *out)
{
    double cr = (wr + 1) / 2.0, cc = (wc + 1) / 2.0;
    double sr = wr / 4.0, sc = wc / 4.0;
    for (int j = 1; j <= wc; j++)
        for (int i = 1; i <= wr; i++) {
            double v = apod == HH_APOD_GAUSS
                           ? exp(-((i - cr) * (i - cr) / (2 * sr * sr) + (j - cc) * (j - cc) / (2 * sc * sc)))
                           : 1.0;
            out[(i - 1) + (size_t)(j - 1) * wr] = (REAL)v;
        }
}

/* fftshift_abs! (src/correlators.jl:282-289): R[mod1(i+n1/2), mod1(j+n2/2)]
 * = abs(C[i, j]); abs of a complex is hypot. */
static void F(orc_fftshift_abs)(const REAL *re, const REAL *im, int n1, int n2, REAL *R)
{
    int s1 = n1 / 2, s2 = n2 / 2;
    for (int j = 0; j < n2; j++)
        for (int i = 0; i < n1; i++) {
            int ip = (i + s1) % n1, jp = (j + s2) % n2;
            R[ip + (size_t)jp * n1] = (REAL)hypot(re[i + (size_t)j * n1], im[i + (size_t)j * n1]);
        }
}

/* overlap_gain! (src/correlators.jl:170-181): autocorrelation of the
 * apodization via FFT in T, fftshifted; wmax/w where w >= wmax/2, else 0. */
static void F(orc_overlap_gain)(F(orc_corr) * c)
{
    int n1 = c->m1, n2 = c->m2;
    size_t np = (size_t)n1 * n2;
    int mx = n1 > n2 ? n1 : n2;
    REAL *re = (REAL *)calloc(np, sizeof(REAL)), *im = (REAL *)calloc(np, sizeof(REAL));
    REAL *lr = (REAL *)malloc(sizeof(REAL) * mx), *li = (REAL *)malloc(sizeof(REAL) * mx);
    REAL *tr = (REAL *)malloc(sizeof(REAL) * mx), *ti = (REAL *)malloc(sizeof(REAL) * mx);
    for (int j = 0; j < c->wc; j++)
        for (int i = 0; i < c->wr; i++) re[i + (size_t)j * n1] = c->apod[i + (size_t)j * c->wr];
    F(orc_fft2)(&c->p1, &c->p2, re, im, 0, lr, li, tr, ti);
    for (size_t k = 0; k < np; k++) {
        re[k] = re[k] * re[k] + im[k] * im[k]; /* abs2 */
        im[k] = 0;
    }
    F(orc_fft2)(&c->p1, &c->p2, re, im, 1, lr, li, tr, ti);
    REAL scale = (REAL)1 / (REAL)(n1 * n2);
    for (size_t k = 0; k < np; k++) {
        re[k] *= scale;
        im[k] *= scale;
    }
    REAL *w = (REAL *)malloc(sizeof(REAL) * np);
    F(orc_fftshift_abs)(re, im, n1, n2, w);
    REAL wmax = w[0];
    for (size_t k = 1; k < np; k++)
        if (w[k] > wmax) wmax = w[k];
    c->gain = (REAL *)malloc(sizeof(REAL) * np);
    for (size_t k = 0; k < np; k++) c->gain[k] = w[k] >= wmax / 2 ? wmax / w[k] : (REAL)0;
    free(re); free(im); free(lr); free(li); free(tr); free(ti); free(w);
}

/* CrossCorrelator{T} / PhaseCorrelator{T} constructors
 * (src/correlators.jl:65-76,110-127): FFT side = search area (twice it when
 * padded); the overlap gain only for padded equal areas. */
static void F(orc_corr_init)(F(orc_corr) * c, const hh_pass *p)
{
    c->wr = p->window_size[0];
    c->wc = p->window_size[1];
    c->sr = p->search_area_size[0];
    c->sc = p->search_area_size[1];
    c->m1 = p->padding ? 2 * c->sr : c->sr;
    c->m2 = p->padding ? 2 * c->sc : c->sc;
    c->method = p->correlation_method;
    c->padding = p->padding;
    F(orc_fftplan_init)(&c->p1, c->m1);
    F(orc_fftplan_init)(&c->p2, c->m2);
    c->apod = (REAL *)malloc(sizeof(REAL) * c->wr * c->wc);
    F(orc_apodization)(c->wr, c->wc, p->apodization, c->apod);
    c->gain = NULL;
    if (p->padding && c->sr == c->wr && c->sc == c->wc) F(orc_overlap_gain)(c);
    c->W = NULL;
    if (p->correlation_method == HH_PHASE) {
        double sigma = (c->m1 < c->m2 ? c->m1 : c->m2) / 8.0;
        c->W = (REAL *)malloc(sizeof(REAL) * c->m1 * c->m2);
        for (int j = 0; j < c->m2; j++)
            for (int i = 0; i < c->m1; i++) {
                /* fftfreq(n, n): 0, 1, ..., then negative frequencies */
                double fi = i < (c->m1 + 1) / 2 ? i : i - c->m1;
                double fj = j < (c->m2 + 1) / 2 ? j : j - c->m2;
                c->W[i + (size_t)j * c->m1] = (REAL)exp(-(fi * fi + fj * fj) / (2 * sigma * sigma));
            }
    }
}

static void F(orc_corr_free)(F(orc_corr) * c)
{
    F(orc_fftplan_free)(&c->p1);
    F(orc_fftplan_free)(&c->p2);
    free(c->apod);
    free(c->gain);
    free(c->W);
}

typedef struct {
    REAL *c1r, *c1i, *c2r, *c2i, *R, *lr, *li, *tr, *ti;
} F(orc_scratch);

static void F(orc_scratch_init)(F(orc_scratch) * s, const F(orc_corr) * c)
{
    size_t np = (size_t)c->m1 * c->m2;
    int mx = c->m1 > c->m2 ? c->m1 : c->m2;
    s->c1r = (REAL *)malloc(sizeof(REAL) * np);
    s->c1i = (REAL *)malloc(sizeof(REAL) * np);
    s->c2r = (REAL *)malloc(sizeof(REAL) * np);
    s->c2i = (REAL *)malloc(sizeof(REAL) * np);
    s->R = (REAL *)malloc(sizeof(REAL) * np);
    s->lr = (REAL *)malloc(sizeof(REAL) * mx);
    s->li = (REAL *)malloc(sizeof(REAL) * mx);
    s->tr = (REAL *)malloc(sizeof(REAL) * mx);
    s->ti = (REAL *)malloc(sizeof(REAL) * mx);
}

static void F(orc_scratch_free)(F(orc_scratch) * s)
{
    free(s->c1r); free(s->c1i); free(s->c2r); free(s->c2i); free(s->R);
    free(s->lr); free(s->li); free(s->tr); free(s->ti);
}

/* correlation_plane! (src/correlators.jl:326-341) for the equal-area,
 * unmasked path of load_windows! (src/correlators.jl:199-207): the mean is
 * a sequential column-major sum in T divided by the pixel count, windows
 * are apod .* (x - mean) at the top-left of the zeroed (padded) buffers;
 * forward FFT both, cross (or phase) spectrum, normalized inverse,
 * fftshift + magnitude, overlap gain. (rs, cs) are 1-based origins. */
/* Window means and apodized windows (load_windows!, src/correlators.jl:
 * 199-227; load_uncertainty_windows!, src/uncertainty.jl:51-79): with a
 * pixel mask (image-sized, nullable) and any masked pixel in the window, the
 * means are over the valid pixels (summed in T, column-major) and masked
 * pixels enter as zero; a clean window takes the unmasked path. */
static int F(orc_window_means)(const REAL *A, const REAL *B, int nr, int rs, int cs, int wr, int wc,
                               const uint8_t *mask, REAL *meanA, REAL *meanB)
{
    int anym = 0;
    if (mask)
        for (int j = 0; j < wc && !anym; j++)
            for (int i = 0; i < wr; i++)
                if (mask[(size_t)(rs - 1 + i) + (size_t)(cs - 1 + j) * nr]) {
                    anym = 1;
                    break;
                }
    REAL sA = 0, sB = 0;
    int n = 0;
    for (int j = 0; j < wc; j++)
        for (int i = 0; i < wr; i++) {
            size_t g = (size_t)(rs - 1 + i) + (size_t)(cs - 1 + j) * nr;
            if (anym && mask[g]) continue;
            sA += A[g];
            sB += B[g];
            n++;
        }
    if (anym) {
        *meanA = n > 0 ? sA / (REAL)n : (REAL)0;
        *meanB = n > 0 ? sB / (REAL)n : (REAL)0;
    } else {
        *meanA = sA / (REAL)(wr * wc);
        *meanB = sB / (REAL)(wr * wc);
    }
    return anym;
}

/* Mean of an h x w block at 1-based (r0, c0) of an image with leading
 * dimension ld, over its unmasked pixels when the block holds a masked one
 * (mask nullable, same coordinates); *anym tells which. */
static REAL F(orc_block_mean)(const REAL *X, int ld, int r0, int c0, int h, int w, const uint8_t *mask, int *anym)
{
    int any = 0;
    if (mask)
        for (int j = 0; j < w && !any; j++)
            for (int i = 0; i < h; i++)
                if (mask[(size_t)(r0 - 1 + i) + (size_t)(c0 - 1 + j) * ld]) {
                    any = 1;
                    break;
                }
    REAL sum = 0;
    int n = 0;
    for (int j = 0; j < w; j++)
        for (int i = 0; i < h; i++) {
            size_t g = (size_t)(r0 - 1 + i) + (size_t)(c0 - 1 + j) * ld;
            if (any && mask[g]) continue;
            sum += X[g];
            n++;
        }
    *anym = any;
    if (any) return n > 0 ? sum / (REAL)n : (REAL)0;
    return sum / (REAL)(h * w);
}

/* correlation_plane! (src/correlators.jl:326-341) of the frame-A window at
 * 1-based (ra, ca) of A (leading dimension lda) and the frame-B search area
 * at (rb, cb) of B (ldb); mask (nullable) in the coordinates of both images
 * (lda = ldb then). Equal areas load as above; an enlarged search area loads
 * as the general path of load_windows! (src/correlators.jl:230-262): the
 * apodized window centred at offset ((sr - wr)/2, (sc - wc)/2) of C1, the
 * mean-subtracted search area (not apodized) in C2, each mean over its own
 * valid pixels; after the magnitude, restrict_search_lags!
 * (src/correlators.jl:347-355) keeps only lags within the margins. */
static void F(orc_plane_at)(const F(orc_corr) * c, F(orc_scratch) * s, const REAL *A, int lda, int ra, int ca,
                            const REAL *B, int ldb, int rb, int cb, const uint8_t *mask)
{
    int wr = c->wr, wc = c->wc, m1 = c->m1;
    size_t np = (size_t)c->m1 * c->m2;
    memset(s->c1r, 0, sizeof(REAL) * np);
    memset(s->c1i, 0, sizeof(REAL) * np);
    memset(s->c2r, 0, sizeof(REAL) * np);
    memset(s->c2i, 0, sizeof(REAL) * np);
    const int equal = c->sr == wr && c->sc == wc;
    if (equal) {
        REAL meanA, meanB;
        const int anym = F(orc_window_means)(A, B, lda, ra, ca, wr, wc, mask, &meanA, &meanB);
        for (int j = 0; j < wc; j++)
            for (int i = 0; i < wr; i++) {
                size_t g = (size_t)(ra - 1 + i) + (size_t)(ca - 1 + j) * lda;
                if (anym && mask[g]) continue; /* masked pixels stay zero */
                REAL a = c->apod[i + (size_t)j * wr];
                s->c1r[i + (size_t)j * m1] = a * (A[g] - meanA);
                s->c2r[i + (size_t)j * m1] = a * (B[g] - meanB);
            }
    } else {
        const int offr = (c->sr - wr) / 2, offc = (c->sc - wc) / 2;
        int anyA = 0, anyB = 0;
        const REAL meanA = F(orc_block_mean)(A, lda, ra, ca, wr, wc, mask, &anyA);
        const REAL meanB = F(orc_block_mean)(B, ldb, rb, cb, c->sr, c->sc, mask, &anyB);
        for (int j = 0; j < wc; j++)
            for (int i = 0; i < wr; i++) {
                size_t g = (size_t)(ra - 1 + i) + (size_t)(ca - 1 + j) * lda;
                if (anyA && mask[g]) continue;
                s->c1r[(offr + i) + (size_t)(offc + j) * m1] = c->apod[i + (size_t)j * wr] * (A[g] - meanA);
            }
        for (int j = 0; j < c->sc; j++)
            for (int i = 0; i < c->sr; i++) {
                size_t g = (size_t)(rb - 1 + i) + (size_t)(cb - 1 + j) * ldb;
                if (anyB && mask[g]) continue;
                s->c2r[i + (size_t)j * m1] = B[g] - meanB;
            }
    }
    F(orc_fft2)(&c->p1, &c->p2, s->c1r, s->c1i, 0, s->lr, s->li, s->tr, s->ti);
    F(orc_fft2)(&c->p1, &c->p2, s->c2r, s->c2i, 0, s->lr, s->li, s->tr, s->ti);
    for (size_t k = 0; k < np; k++) {
        /* conj(C1) * C2 (spectrum!, src/correlators.jl:266-279) */
        REAL ar = s->c1r[k], ai = s->c1i[k], br = s->c2r[k], bi = s->c2i[k];
        REAL pr = ar * br + ai * bi;
        REAL pi = ar * bi - ai * br;
        if (c->method == HH_PHASE) {
            REAL den = (REAL)hypot(pr, pi) + REAL_EPS;
            REAL wv = c->W[k];
            pr = wv * pr / den;
            pi = wv * pi / den;
        }
        s->c1r[k] = pr;
        s->c1i[k] = pi;
    }
    F(orc_fft2)(&c->p1, &c->p2, s->c1r, s->c1i, 1, s->lr, s->li, s->tr, s->ti);
    REAL scale = (REAL)1 / (REAL)np; /* inv(plan) normalization */
    for (size_t k = 0; k < np; k++) {
        s->c1r[k] *= scale;
        s->c1i[k] *= scale;
    }
    F(orc_fftshift_abs)(s->c1r, s->c1i, c->m1, c->m2, s->R);
    if (c->gain)
        for (size_t k = 0; k < np; k++) s->R[k] *= c->gain[k];
    if (!equal) {
        const int mr = (c->sr - wr) / 2, mc = (c->sc - wc) / 2, cr = c->m1 / 2, cc = c->m2 / 2;
        for (int j = 0; j < c->m2; j++)
            for (int i = 0; i < c->m1; i++)
                if (!(abs(i - cr) <= mr && abs(j - cc) <= mc)) s->R[i + (size_t)j * c->m1] = 0;
    }
}

/* The plane of job (rs, cs) (window origin, 1-based) of an image pair with
 * leading dimension nr: frame-B search area at (rs - mr, cs - mc). */
static void F(orc_plane)(const F(orc_corr) * c, F(orc_scratch) * s, const REAL *A, const REAL *B, int nr,
                         int rs, int cs, const uint8_t *mask)
{
    F(orc_plane_at)(c, s, A, nr, rs, cs, B, nr, rs - (c->sr - c->wr) / 2, cs - (c->sc - c->wc) / 2, mask);
}

/* ======================================================================
 * Peak analysis (src/localmax.jl, src/correlators.jl:390-564,
 * src/quality.jl:11-74)
 * ==================================================================== */

/* is_local_maximum (src/localmax.jl:3-19); r, c 1-based. */
int F(orc_is_local_maximum)(const REAL *A, int nr, int nc, int r, int c)
{
    REAL I0 = A[(r - 1) + (size_t)(c - 1) * nr];
    if (isnan(I0)) return 0;
    long lin0 = r + (long)(c - 1) * nr;
    int jlo = c - 1 > 1 ? c - 1 : 1, jhi = c + 1 < nc ? c + 1 : nc;
    int ilo = r - 1 > 1 ? r - 1 : 1, ihi = r + 1 < nr ? r + 1 : nr;
    for (int jj = jlo; jj <= jhi; jj++)
        for (int ii = ilo; ii <= ihi; ii++) {
            if (ii == r && jj == c) continue;
            REAL In = A[(ii - 1) + (size_t)(jj - 1) * nr];
            if (ii + (long)(jj - 1) * nr < lin0) {
                if (In > I0) return 0;
            } else {
                if (In >= I0) return 0;
            }
        }
    return 1;
}

/* find_peaks_regionalmax! (src/correlators.jl:426-454). locs: 2 x k
 * (row, col) 1-based. Returns found. */
int F(orc_find_peaks_regionalmax)(const REAL *R, int nr, int nc, int k, REAL *vals, int32_t *locs)
{
    int nfound = 0;
    for (int j = 1; j <= nc; j++)
        for (int i = 1; i <= nr; i++) {
            if (!F(orc_is_local_maximum)(R, nr, nc, i, j)) continue;
            REAL val = R[(i - 1) + (size_t)(j - 1) * nr];
            if (!(nfound < k || val > vals[nfound - 1])) continue;
            int pos = nfound;
            while (pos >= 1 && val > vals[pos - 1]) pos--;
            int ins = pos + 1;
            if (ins > k) continue;
            if (nfound < k) nfound++;
            for (int m = nfound; m >= ins + 1; m--) {
                vals[m - 1] = vals[m - 2];
                locs[2 * (m - 1)] = locs[2 * (m - 2)];
                locs[2 * (m - 1) + 1] = locs[2 * (m - 2) + 1];
            }
            vals[ins - 1] = val;
            locs[2 * (ins - 1)] = i;
            locs[2 * (ins - 1) + 1] = j;
        }
    if (nfound == 0) return 0;
    int found = 1;
    while (found < nfound && vals[found] > 0) found++;
    return found;
}

/* find_peaks_exclusion! (src/correlators.jl:390-421). */
int F(orc_find_peaks_exclusion)(const REAL *R, int nr, int nc, int k, int radius, REAL *vals, int32_t *locs)
{
    int found = 0;
    for (int t = 0; t < k; t++) {
        REAL best = (REAL)-INFINITY;
        int bi = 0, bj = 0;
        for (int j = 1; j <= nc; j++)
            for (int i = 1; i <= nr; i++) {
                int excluded = 0;
                for (int p = 0; p < found; p++)
                    if (abs(i - locs[2 * p]) <= radius && abs(j - locs[2 * p + 1]) <= radius) {
                        excluded = 1;
                        break;
                    }
                if (excluded) continue;
                REAL v = R[(i - 1) + (size_t)(j - 1) * nr];
                if (v > best) {
                    best = v;
                    bi = i;
                    bj = j;
                }
            }
        if (bi == 0) break;
        if (found > 0 && best <= 0) break;
        vals[found] = best;
        locs[2 * found] = bi;
        locs[2 * found + 1] = bj;
        found++;
    }
    return found;
}

/* subpixel_gauss3 (src/correlators.jl:499-520); logs in T. out: (row, col). */
void F(orc_gauss3)(const REAL *R, int nr, int nc, int i, int j, REAL *out)
{
    REAL di = 0, dj = 0;
    REAL I0 = R[(i - 1) + (size_t)(j - 1) * nr];
    if (1 < i && i < nr) {
        REAL Im = R[(i - 2) + (size_t)(j - 1) * nr], Ip = R[i + (size_t)(j - 1) * nr];
        if (Im > 0 && I0 > 0 && Ip > 0) {
            REAL denom = RLOG(Im) - 2 * RLOG(I0) + RLOG(Ip);
            di = denom != 0 ? (RLOG(Im) - RLOG(Ip)) / (2 * denom) : (REAL)0;
        }
    }
    if (1 < j && j < nc) {
        REAL Im = R[(i - 1) + (size_t)(j - 2) * nr], Ip = R[(i - 1) + (size_t)j * nr];
        if (Im > 0 && I0 > 0 && Ip > 0) {
            REAL denom = RLOG(Im) - 2 * RLOG(I0) + RLOG(Ip);
            dj = denom != 0 ? (RLOG(Im) - RLOG(Ip)) / (2 * denom) : (REAL)0;
        }
    }
    out[0] = (REAL)i + di;
    out[1] = (REAL)j + dj;
}

/* subpixel_gauss9 (src/correlators.jl:534-564). */
void F(orc_gauss9)(const REAL *R, int nr, int nc, int pr, int pc, REAL *out)
{
    if (!(1 < pr && pr < nr && 1 < pc && pc < nc)) {
        F(orc_gauss3)(R, nr, nc, pr, pc, out);
        return;
    }
    REAL sL = 0, sxL = 0, syL = 0, sxxL = 0, syyL = 0, sxyL = 0;
    for (int dj = -1; dj <= 1; dj++)
        for (int di = -1; di <= 1; di++) {
            REAL val = R[(pr + di - 1) + (size_t)(pc + dj - 1) * nr];
            if (!(val > 0)) {
                F(orc_gauss3)(R, nr, nc, pr, pc, out);
                return;
            }
            REAL L = RLOG(val);
            sL += L;
            sxL += dj * L;
            syL += di * L;
            sxxL += dj * dj * L;
            syyL += di * di * L;
            sxyL += di * dj * L;
        }
    REAL a1 = sxL / 6, a2 = syL / 6, a3 = sxxL / 2 - sL / 3, a4 = sxyL / 4, a5 = syyL / 2 - sL / 3;
    REAL D = 4 * a3 * a5 - a4 * a4;
    if (!(a3 < 0 && D > 0)) {
        F(orc_gauss3)(R, nr, nc, pr, pc, out);
        return;
    }
    REAL dx = (a4 * a2 - 2 * a5 * a1) / D;
    REAL dy = (a4 * a1 - 2 * a3 * a2) / D;
    if (!(fabs((double)dx) <= 1 && fabs((double)dy) <= 1)) {
        F(orc_gauss3)(R, nr, nc, pr, pc, out);
        return;
    }
    out[0] = (REAL)pr + dy;
    out[1] = (REAL)pc + dx;
}

/* calculate_correlation_moment, neighborhood 3 (src/quality.jl:49-74):
 * Float64 sums; (c - pc)^2 is computed in T, then promoted. */
REAL F(orc_moment)(const REAL *R, int nr, int nc, REAL pr, REAL pc)
{
    if (!(isfinite(pr) && isfinite(pc))) return (REAL)NAN;
    int r0 = (int)nearbyint((double)pr), c0 = (int)nearbyint((double)pc); /* round half even */
    int rlo = r0 - 1 > 1 ? r0 - 1 : 1, rhi = r0 + 1 < nr ? r0 + 1 : nr;
    int clo = c0 - 1 > 1 ? c0 - 1 : 1, chi = c0 + 1 < nc ? c0 + 1 : nc;
    if (rlo > rhi || clo > chi) return (REAL)NAN;
    double sumC = 0, sdx = 0, sdy = 0;
    for (int c = clo; c <= chi; c++)
        for (int r = rlo; r <= rhi; r++) {
            double w = (double)R[(r - 1) + (size_t)(c - 1) * nr];
            if (!(w > 0.0)) w = 0.0; /* max(Float64(R), 0.0) */
            REAL dxr = (REAL)c - pc, dyr = (REAL)r - pr;
            sumC += w;
            sdx += (double)(dxr * dxr) * w;
            sdy += (double)(dyr * dyr) * w;
        }
    if (sumC < 1e-9) return (REAL)NAN;
    return (REAL)sqrt((sdx + sdy) / sumC);
}

/* analyze_plane! (src/correlators.jl:459-468) plus the alternatives of
 * process_windows! (src/pipeline.jl:1009-1017). out[0..3] = du, dv, ratio,
 * moment; alt (n_peaks-1 pairs du, dv) written only for found peaks.
 * vals/locs scratch of K = max(n_peaks, 2). Returns found. */
int F(orc_analyze_plane)(const REAL *R, int nr, int nc, const hh_pass *p, REAL *vals, int32_t *locs, REAL *out,
                         REAL *alt_du, REAL *alt_dv)
{
    int K = p->n_peaks > 2 ? p->n_peaks : 2;
    int found = p->peak_finder == HH_EXCLUSION ? F(orc_find_peaks_exclusion)(R, nr, nc, K, 2, vals, locs)
                                               : F(orc_find_peaks_regionalmax)(R, nr, nc, K, vals, locs);
    int cr = nr / 2 + 1, cc = nc / 2 + 1;
    if (found == 0) {
        /* degenerate all-NaN plane: the reference reads uninitialized scratch;
         * any well-defined value is acceptable (src/ka_backend.jl:659-665) */
        out[0] = out[1] = out[2] = out[3] = (REAL)NAN;
        return 0;
    }
    REAL ref[2];
    if (p->subpixel_method == HH_GAUSS9)
        F(orc_gauss9)(R, nr, nc, locs[0], locs[1], ref);
    else if (p->subpixel_method == HH_SUBPIX_NONE) {
        ref[0] = (REAL)locs[0];
        ref[1] = (REAL)locs[1];
    } else
        F(orc_gauss3)(R, nr, nc, locs[0], locs[1], ref);
    out[0] = ref[1] - (REAL)cc;
    out[1] = ref[0] - (REAL)cr;
    out[2] = found >= 2 ? vals[0] / vals[1] : (vals[0] > 0 ? (REAL)INFINITY : (REAL)NAN);
    out[3] = F(orc_moment)(R, nr, nc, ref[0], ref[1]);
    int mlim = found < p->n_peaks ? found : p->n_peaks;
    for (int m = 2; m <= mlim; m++) {
        REAL a[2];
        F(orc_gauss3)(R, nr, nc, locs[2 * (m - 1)], locs[2 * (m - 1) + 1], a);
        alt_du[m - 2] = a[1] - (REAL)cc;
        alt_dv[m - 2] = a[0] - (REAL)cr;
    }
    return found;
}

/* ======================================================================
 * Wieneke (2015) uncertainty (src/uncertainty.jl:16-179)
 * ==================================================================== */

/* uq_component! (src/uncertainty.jl:104-151). A, B: nr x nc with the shift
 * along columns; element (r, c) at A[r*rs + c*cs] (transposed views pass
 * swapped strides). dC/dCs scratch nr x (nc-1) in T. */
static void F(orc_uq_component)(double *s, const REAL *A, const REAL *B, int nr, int nc, int rs, int cs, REAL *dC,
                                REAL *dCs, int off[ORC_UQ_NOFF][2])
{
    int m = nc - 1;
    double C0 = 0, Cp = 0, Cm = 0;
    for (int c = 0; c < m; c++)
        for (int r = 0; r < nr; r++) {
            REAL a0 = A[r * rs + c * cs], a1 = A[r * rs + (c + 1) * cs];
            REAL b0 = B[r * rs + c * cs], b1 = B[r * rs + (c + 1) * cs];
            REAL p00 = a0 * b0, p01 = a0 * b1, p10 = a1 * b0;
            C0 += (double)p00;
            Cp += (double)p01;
            Cm += (double)p10;
            dC[r + c * nr] = p01 - p10;
        }
    s[0] += C0;
    s[1] += Cp;
    s[2] += Cm;
    REAL quarter = (REAL)1 / 4;
    for (int c = 0; c < m; c++)
        for (int r = 0; r < nr; r++) {
            int cl = c - 1 > 0 ? c - 1 : 0, cr = c + 1 < m - 1 ? c + 1 : m - 1;
            dCs[r + c * nr] = quarter * (dC[r + cl * nr] + 2 * dC[r + c * nr] + dC[r + cr * nr]);
        }
    double mu = 0;
    for (int c = 0; c < m; c++)
        for (int r = 0; r < nr; r++) mu += (double)dCs[r + c * nr];
    mu /= (double)(nr * m);
    double S00 = 0;
    for (int c = 0; c < m; c++)
        for (int r = 0; r < nr; r++) {
            double d = (double)dCs[r + c * nr] - mu;
            S00 += d * d;
        }
    s[3] += S00;
    for (int k = 0; k < ORC_UQ_NOFF; k++) {
        int dr = off[k][0], dc = off[k][1];
        double Sk = 0;
        int c_lo = 1 - dc > 1 ? 1 - dc : 1, c_hi = m - dc < m ? m - dc : m;
        int r_lo = 1 - dr > 1 ? 1 - dr : 1, r_hi = nr - dr < nr ? nr - dr : nr;
        for (int c = c_lo; c <= c_hi; c++)
            for (int r = r_lo; r <= r_hi; r++)
                Sk += ((double)dCs[(r - 1) + (c - 1) * nr] - mu) *
                      ((double)dCs[(r - 1 + dr) + (c - 1 + dc) * nr] - mu);
        s[4 + k] += Sk;
    }
}

/* accumulate_uncertainty! with load_uncertainty_windows!
 * (src/uncertainty.jl:51-96); mask nullable (image-sized). stats: 2 x 44,
 * row-major per component. */
void F(orc_uq_accumulate)(double *stats, const REAL *A, const REAL *B, int nr, int rs, int cs, int wr, int wc,
                          const REAL *apod, REAL *wA, REAL *wB, REAL *dC, REAL *dCs, const uint8_t *mask)
{
    int off[ORC_UQ_NOFF][2], rings[ORC_UQ_MAXOFF + 1];
    orc_uq_offsets(off, rings);
    REAL meanA, meanB;
    const int anym = F(orc_window_means)(A, B, nr, rs, cs, wr, wc, mask, &meanA, &meanB);
    for (int j = 0; j < wc; j++)
        for (int i = 0; i < wr; i++) {
            size_t g = (size_t)(rs - 1 + i) + (size_t)(cs - 1 + j) * nr;
            if (anym && mask[g]) {
                wA[i + j * wr] = wB[i + j * wr] = (REAL)0;
                continue;
            }
            REAL a = apod[i + j * wr];
            wA[i + j * wr] = a * (A[g] - meanA);
            wB[i + j * wr] = a * (B[g] - meanB);
        }
    /* u: shift along columns; v: transposed views (rows become columns) */
    F(orc_uq_component)(stats, wA, wB, wr, wc, 1, wr, dC, dCs, off);
    F(orc_uq_component)(stats + HH_UQ_NSTATS, wA, wB, wc, wr, wr, 1, dC, dCs, off);
}

/* finalize_uncertainty (src/uncertainty.jl:162-179). */
REAL F(orc_uq_finalize)(const double *s)
{
    int off[ORC_UQ_NOFF][2], rings[ORC_UQ_MAXOFF + 1];
    orc_uq_offsets(off, rings);
    double C0 = s[0], Cp = s[1], Cm = s[2], S00 = s[3];
    double s2 = S00;
    for (int ring = 0; ring < ORC_UQ_MAXOFF; ring++) {
        double mx = -INFINITY;
        for (int k = rings[ring]; k < rings[ring + 1]; k++)
            if (s[4 + k] > mx) mx = s[4 + k];
        if (mx < 0.05 * S00) break;
        for (int k = rings[ring]; k < rings[ring + 1]; k++) s2 += 2 * s[4 + k];
    }
    double sdc = sqrt(s2 > 0 ? s2 : 0.0);
    double Cpm = (Cp + Cm) / 2;
    double lo = Cpm - sdc / 2, hi = Cpm + sdc / 2;
    if (!(C0 > 0 && lo > 0)) return (REAL)NAN;
    double denom = 2 * log(C0) - log(hi) - log(lo);
    if (!(denom > 0)) return (REAL)NAN;
    return (REAL)((log(hi) - log(lo)) / (2 * denom));
}

/* ======================================================================
 * Statistics helpers: median!/middle and type-7 quantile! semantics
 * (Statistics.jl, used at src/quality.jl:107-108,280-287,330-335 and
 * src/pipeline.jl:865)
 * ==================================================================== */
static int F(orc_cmp)(const void *a, const void *b)
{
    REAL x = *(const REAL *)a, y = *(const REAL *)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

/* median! in T: NaN if any NaN; odd n the middle element; even n
 * middle(a, b) = a/2 + b/2. Sorts v. */
REAL F(orc_median)(REAL *v, int n)
{
    for (int i = 0; i < n; i++)
        if (isnan(v[i])) return (REAL)NAN;
    qsort(v, n, sizeof(REAL), F(orc_cmp));
    if (n & 1) return v[n / 2];
    return v[n / 2 - 1] / 2 + v[n / 2] / 2;
}

/* quantile!(v, p) (type 7): aleph = n p + (1 - p), j = clamp(trunc(aleph),
 * 1, n-1), gamma = aleph - j; a + gamma (b - a) if both finite else
 * (1 - gamma) a + gamma b. Sorts v. */
REAL F(orc_quantile)(REAL *v, int n, double p)
{
    qsort(v, n, sizeof(REAL), F(orc_cmp));
    double aleph = n * p + (1.0 - p);
    int j = (int)aleph;
    if (j > n - 1) j = n - 1;
    if (j < 1) j = 1;
    REAL gamma = (REAL)(aleph - j);
    if (gamma < 0) gamma = 0;
    if (gamma > 1) gamma = 1;
    REAL a, b;
    if (n == 1) {
        a = v[0];
        b = v[0];
    } else {
        a = v[j - 1];
        b = v[j];
    }
    if (isfinite(a) && isfinite(b)) return a + gamma * (b - a);
    return (1 - gamma) * a + gamma * b;
}

/* median(::Vector{Float64}) */
static double F(orc_median64)(double *v, int n)
{
    for (int i = 0; i < n; i++)
        if (isnan(v[i])) return NAN;
    qsort(v, n, sizeof(double), orc_cmpd);
    if (n & 1) return v[n / 2];
    return v[n / 2 - 1] / 2 + v[n / 2] / 2;
}

/* ======================================================================
 * Validation and replacement (src/quality.jl:85-136,244-351,467-503;
 * src/pipeline.jl:617-645)
 * ==================================================================== */

/* universal_outlier_detection (src/quality.jl:244-293); ORs into flags.
 * exclude nullable. Medians and residuals in T; the normalized residual is
 * T / (Float64(median) + epsilon) in Float64. */
void F(orc_uod)(const REAL *u, const REAL *v, int nr, int nc, double thr, int nb, double eps,
                const uint8_t *exclude, uint8_t *flags)
{
    int maxn = (2 * nb + 1) * (2 * nb + 1) - 1;
    REAL *nu = (REAL *)malloc(sizeof(REAL) * maxn), *nv = (REAL *)malloc(sizeof(REAL) * maxn);
    REAL *ru = (REAL *)malloc(sizeof(REAL) * maxn), *rv = (REAL *)malloc(sizeof(REAL) * maxn);
    for (int c = 0; c < nc; c++)
        for (int r = 0; r < nr; r++) {
            if (exclude && exclude[r + (size_t)c * nr]) continue;
            int n = 0;
            int c2lo = c - nb > 0 ? c - nb : 0, c2hi = c + nb < nc - 1 ? c + nb : nc - 1;
            int r2lo = r - nb > 0 ? r - nb : 0, r2hi = r + nb < nr - 1 ? r + nb : nr - 1;
            for (int c2 = c2lo; c2 <= c2hi; c2++)
                for (int r2 = r2lo; r2 <= r2hi; r2++) {
                    if (r2 == r && c2 == c) continue;
                    if (exclude && exclude[r2 + (size_t)c2 * nr]) continue;
                    nu[n] = u[r2 + (size_t)c2 * nr];
                    nv[n] = v[r2 + (size_t)c2 * nr];
                    n++;
                }
            if (n == 0) continue;
            REAL umed = F(orc_median)(nu, n), vmed = F(orc_median)(nv, n);
            for (int i = 0; i < n; i++) {
                ru[i] = (REAL)fabs((double)(nu[i] - umed));
                rv[i] = (REAL)fabs((double)(nv[i] - vmed));
            }
            REAL du = (REAL)fabs((double)(u[r + (size_t)c * nr] - umed));
            REAL dv = (REAL)fabs((double)(v[r + (size_t)c * nr] - vmed));
            double un = (double)du / ((double)F(orc_median)(ru, n) + eps);
            double vn = (double)dv / ((double)F(orc_median)(rv, n) + eps);
            if (un > thr || vn > thr) flags[r + (size_t)c * nr] = 1;
        }
    free(nu); free(nv); free(ru); free(rv);
}

/* substitute_alternatives! (src/quality.jl:307-351); alt arrays
 * nr x nc x m. Returns the number substituted. */
int F(orc_substitute)(REAL *u, REAL *v, uint8_t *outliers, const uint8_t *mask, int nr, int nc, const REAL *alt_u,
                      const REAL *alt_v, int m, int er, double thr)
{
    size_t np = (size_t)nr * nc;
    int any = 0;
    for (size_t k = 0; k < np; k++) any |= outliers[k];
    if (!any) return 0;
    REAL *uref = (REAL *)malloc(sizeof(REAL) * np), *vref = (REAL *)malloc(sizeof(REAL) * np);
    uint8_t *valid = (uint8_t *)malloc(np);
    memcpy(uref, u, sizeof(REAL) * np);
    memcpy(vref, v, sizeof(REAL) * np);
    for (size_t k = 0; k < np; k++) valid[k] = !(outliers[k] || (mask && mask[k]));
    int maxn = (2 * er + 1) * (2 * er + 1);
    double *bu = (double *)malloc(sizeof(double) * maxn), *bv = (double *)malloc(sizeof(double) * maxn);
    int nsub = 0;
    for (int c = 0; c < nc; c++)
        for (int r = 0; r < nr; r++) {
            size_t k0 = r + (size_t)c * nr;
            if (!(outliers[k0] && !(mask && mask[k0]))) continue;
            int n = 0;
            for (int c2 = (c - er > 0 ? c - er : 0); c2 <= (c + er < nc - 1 ? c + er : nc - 1); c2++)
                for (int r2 = (r - er > 0 ? r - er : 0); r2 <= (r + er < nr - 1 ? r + er : nr - 1); r2++) {
                    if (r2 == r && c2 == c) continue;
                    if (!valid[r2 + (size_t)c2 * nr]) continue;
                    bu[n] = (double)uref[r2 + (size_t)c2 * nr];
                    bv[n] = (double)vref[r2 + (size_t)c2 * nr];
                    n++;
                }
            if (n < 3) continue;
            double mu = F(orc_median64)(bu, n), mv = F(orc_median64)(bv, n);
            for (int i = 0; i < n; i++) {
                bu[i] = fabs(bu[i] - mu);
                bv[i] = fabs(bv[i] - mv);
            }
            double madu = F(orc_median64)(bu, n), madv = F(orc_median64)(bv, n);
            for (int q = 0; q < m; q++) {
                REAL au = alt_u[k0 + q * np], av = alt_v[k0 + q * np];
                if (!(isfinite(au) && isfinite(av))) break;
                if (fabs((double)au - mu) / (madu + 0.1) <= thr && fabs((double)av - mv) / (madv + 0.1) <= thr) {
                    u[k0] = au;
                    v[k0] = av;
                    outliers[k0] = 0;
                    nsub++;
                    break;
                }
            }
        }
    free(uref); free(vref); free(valid); free(bu); free(bv);
    return nsub;
}

/* replace_vectors! (src/quality.jl:85-114): growing squares until >= 3
 * valid, Float64 median stored as T, from a snapshot. */
void F(orc_replace)(REAL *u, REAL *v, const uint8_t *invalid, int nr, int nc)
{
    size_t np = (size_t)nr * nc;
    int any = 0;
    for (size_t k = 0; k < np; k++) any |= invalid[k];
    if (!any) return;
    REAL *uref = (REAL *)malloc(sizeof(REAL) * np), *vref = (REAL *)malloc(sizeof(REAL) * np);
    memcpy(uref, u, sizeof(REAL) * np);
    memcpy(vref, v, sizeof(REAL) * np);
    double *bu = (double *)malloc(sizeof(double) * np), *bv = (double *)malloc(sizeof(double) * np);
    int maxr = (nr > nc ? nr : nc) - 1;
    for (int c = 0; c < nc; c++)
        for (int r = 0; r < nr; r++) {
            if (!invalid[r + (size_t)c * nr]) continue;
            for (int rad = 1; rad <= maxr; rad++) {
                int n = 0;
                for (int c2 = (c - rad > 0 ? c - rad : 0); c2 <= (c + rad < nc - 1 ? c + rad : nc - 1); c2++)
                    for (int r2 = (r - rad > 0 ? r - rad : 0); r2 <= (r + rad < nr - 1 ? r + rad : nr - 1); r2++) {
                        if (invalid[r2 + (size_t)c2 * nr]) continue;
                        bu[n] = (double)uref[r2 + (size_t)c2 * nr];
                        bv[n] = (double)vref[r2 + (size_t)c2 * nr];
                        n++;
                    }
                if (n >= 3) {
                    u[r + (size_t)c * nr] = (REAL)F(orc_median64)(bu, n);
                    v[r + (size_t)c * nr] = (REAL)F(orc_median64)(bv, n);
                    break;
                }
            }
        }
    free(uref); free(vref); free(bu); free(bv);
}

/* smooth_field (src/quality.jl:123-136): separable [1 2 1]/4, replicated. */
void F(orc_smooth)(const REAL *f, int nr, int nc, REAL *out)
{
    REAL q = (REAL)1 / 4;
    REAL *tmp = (REAL *)malloc(sizeof(REAL) * nr * nc);
    for (int c = 0; c < nc; c++)
        for (int r = 0; r < nr; r++) {
            int cl = c - 1 > 0 ? c - 1 : 0, cr = c + 1 < nc - 1 ? c + 1 : nc - 1;
            tmp[r + (size_t)c * nr] = q * (f[r + (size_t)cl * nr] + 2 * f[r + (size_t)c * nr] + f[r + (size_t)cr * nr]);
        }
    for (int c = 0; c < nc; c++)
        for (int r = 0; r < nr; r++) {
            int rl = r - 1 > 0 ? r - 1 : 0, rr = r + 1 < nr - 1 ? r + 1 : nr - 1;
            out[r + (size_t)c * nr] =
                q * (tmp[rl + (size_t)c * nr] + 2 * tmp[r + (size_t)c * nr] + tmp[rr + (size_t)c * nr]);
        }
    free(tmp);
}

/* validate_and_replace! (src/pipeline.jl:617-645) with the validators of
 * src/quality.jl:479-503. alt arrays nullable (ny x nx x (n_peaks-1)). */
void F(orc_validate_and_replace)(const hh_pass *p, int nr, int nc, int force_replace, REAL *u, REAL *v,
                                 const REAL *pr, const REAL *cm, const REAL *alt_u, const REAL *alt_v,
                                 const uint8_t *mask, uint8_t *outliers)
{
    size_t np = (size_t)nr * nc;
    int anymask = 0;
    if (mask)
        for (size_t k = 0; k < np; k++) anymask |= mask[k];
    memset(outliers, 0, np);
    if (p->uod_enable)
        F(orc_uod)(u, v, nr, nc, p->uod_threshold, p->uod_neighborhood, 0.1, anymask ? mask : NULL, outliers);
    if (p->min_peak_ratio > 1)
        for (size_t k = 0; k < np; k++)
            if (!((double)pr[k] >= p->min_peak_ratio)) outliers[k] = 1;
    for (int q = 0; q < p->n_validators; q++) {
        const hh_validator *vd = &p->validators[q];
        if (vd->kind == HH_V_PEAK_RATIO) {
            for (size_t k = 0; k < np; k++)
                if (!((double)pr[k] >= vd->a)) outliers[k] = 1;
        } else if (vd->kind == HH_V_MOMENT) {
            for (size_t k = 0; k < np; k++)
                if (!((double)cm[k] <= vd->a)) outliers[k] = 1;
        } else if (vd->kind == HH_V_VELMAG) {
            for (size_t k = 0; k < np; k++) {
                double mag = (double)(REAL)hypot((double)u[k], (double)v[k]);
                if (!(vd->a <= mag && mag <= vd->b)) outliers[k] = 1;
            }
        } else if (vd->kind == HH_V_UOD) {
            F(orc_uod)(u, v, nr, nc, vd->a, vd->neighborhood, vd->b, anymask ? mask : NULL, outliers);
        }
    }
    if (mask)
        for (size_t k = 0; k < np; k++) outliers[k] &= !mask[k];
    if (alt_u && p->n_peaks > 1)
        F(orc_substitute)(u, v, outliers, mask, nr, nc, alt_u, alt_v, p->n_peaks - 1, p->uod_neighborhood,
                          p->uod_threshold);
    if (force_replace || p->replace_outliers) {
        uint8_t *inv = (uint8_t *)malloc(np);
        int any = 0;
        for (size_t k = 0; k < np; k++) {
            inv[k] = outliers[k] | (mask ? mask[k] : 0);
            any |= inv[k];
        }
        if (any) {
            F(orc_replace)(u, v, inv, nr, nc);
            if (mask)
                for (size_t k = 0; k < np; k++)
                    if (mask[k]) u[k] = v[k] = (REAL)NAN;
        }
        free(inv);
    }
}

/* ======================================================================
 * Deformation (src/pipeline.jl:561-600,909-965; src/ka_backend.jl:257-312)
 * ==================================================================== */

/* Natural cubic B-spline coefficients of one line: the (n+2)-unknown system
 * of Interpolations' BSpline(Cubic(Line(OnGrid()))): rows k = 1..n:
 * (c[k-1] + 4 c[k] + c[k+1])/6 = y[k]; boundary rows c[0] - 2c[1] + c[2] = 0
 * and c[n-1] - 2c[n] + c[n+1] = 0. These give c[1] = y[1], c[n] = y[n];
 * the interior is a [1 4 1] tridiagonal solve (Thomas). c has n+2 entries. */
static void F(orc_spline_line)(const REAL *y, int ys, int n, REAL *c, REAL *cp, REAL *dp)
{
    c[1] = y[0];
    c[n] = y[(size_t)(n - 1) * ys];
    int m = n - 2;
    if (m >= 1) {
        for (int t = 0; t < m; t++) {
            REAL rhs = 6 * y[(size_t)(t + 1) * ys];
            if (t == 0) rhs -= c[1];
            if (t == m - 1) rhs -= c[n];
            if (t == 0) {
                cp[0] = (REAL)1 / 4;
                dp[0] = rhs / 4;
            } else {
                REAL den = 4 - cp[t - 1];
                cp[t] = 1 / den;
                dp[t] = (rhs - dp[t - 1]) / den;
            }
        }
        c[m + 1] = dp[m - 1];
        for (int t = m - 2; t >= 0; t--) c[t + 2] = dp[t] - cp[t] * c[t + 3];
    }
    c[0] = 2 * c[1] - c[2];
    c[n + 1] = 2 * c[n] - c[n - 1];
}

/* image_interpolant (src/pipeline.jl:909-915): padded (nr+2) x (nc+2)
 * tensor-product coefficients, first along dim 1 then along dim 2. */
void F(orc_prefilter)(const REAL *img, int nr, int nc, REAL *coef)
{
    int N1 = nr + 2, N2 = nc + 2;
    int mx = (N1 > N2 ? N1 : N2) + 2;
    REAL *tmp = (REAL *)malloc(sizeof(REAL) * (size_t)N1 * nc);
    REAL *cp = (REAL *)malloc(sizeof(REAL) * mx), *dp = (REAL *)malloc(sizeof(REAL) * mx);
    REAL *line = (REAL *)malloc(sizeof(REAL) * mx), *out = (REAL *)malloc(sizeof(REAL) * mx);
    for (int j = 0; j < nc; j++) F(orc_spline_line)(img + (size_t)j * nr, 1, nr, tmp + (size_t)j * N1, cp, dp);
    for (int i = 0; i < N1; i++) {
        for (int j = 0; j < nc; j++) line[j] = tmp[i + (size_t)j * N1];
        F(orc_spline_line)(line, 1, nc, out, cp, dp);
        for (int j = 0; j < N2; j++) coef[i + (size_t)j * N1] = out[j];
    }
    free(tmp); free(cp); free(dp); free(line); free(out);
}

/* _pk_lin_axis (src/ka_backend.jl:257-263): 0-based interval index and
 * fraction over uniform knots q0 + i*qstep, Flat() clamping. */
static void F(orc_lin_axis)(REAL q0, REAL qstep, int n, REAL q, int *i, REAL *l)
{
    if (n == 1) {
        *i = 0;
        *l = 0;
        return;
    }
    REAL t = (q - q0) / qstep;
    if (t < 0) t = 0;
    if (t > (REAL)(n - 1)) t = (REAL)(n - 1);
    int ii = (int)floor((double)t);
    if (ii > n - 2) ii = n - 2;
    *i = ii;
    *l = t - (REAL)ii;
}

/* Bilinear (Gridded(Linear()), Flat()) value of predictor grid p
 * (gny x gnx) at (y, x): the evaluation order of _ka_deform!. */
static REAL F(orc_bilinear)(const REAL *p, int gny, int gnx, REAL y0, REAL ysp, REAL x0, REAL xsp, REAL y, REAL x)
{
    int iy, ix;
    REAL ly, lx;
    F(orc_lin_axis)(y0, ysp, gny, y, &iy, &ly);
    F(orc_lin_axis)(x0, xsp, gnx, x, &ix, &lx);
    int iy2 = gny == 1 ? iy : iy + 1, ix2 = gnx == 1 ? ix : ix + 1;
    REAL a = p[iy + ix * gny] + lx * (p[iy + ix2 * gny] - p[iy + ix * gny]);
    REAL b = p[iy2 + ix * gny] + lx * (p[iy2 + ix2 * gny] - p[iy2 + ix * gny]);
    return a + ly * (b - a);
}

/* _pk_cubic (src/ka_backend.jl:276-291): zero outside [1, n]; C is the
 * padded (nr+2) x (nc+2) coefficient array. */
REAL F(orc_cubic)(const REAL *C, int nr, int nc, REAL yy, REAL xx)
{
    if (yy < 1 || yy > (REAL)nr || xx < 1 || xx > (REAL)nc) return 0;
    int iy = (int)floor((double)yy);
    if (iy > nr - 1) iy = nr - 1;
    int ix = (int)floor((double)xx);
    if (ix > nc - 1) ix = nc - 1;
    REAL dy = yy - (REAL)iy, dx = xx - (REAL)ix;
    REAL wy[4] = {(1 - dy) * (1 - dy) * (1 - dy) / 6, (4 - 6 * dy * dy + 3 * dy * dy * dy) / 6,
                  (1 + 3 * dy + 3 * dy * dy - 3 * dy * dy * dy) / 6, dy * dy * dy / 6};
    REAL wx[4] = {(1 - dx) * (1 - dx) * (1 - dx) / 6, (4 - 6 * dx * dx + 3 * dx * dx * dx) / 6,
                  (1 + 3 * dx + 3 * dx * dx - 3 * dx * dx * dx) / 6, dx * dx * dx / 6};
    int N1 = nr + 2;
    REAL s = 0;
    for (int n = 0; n < 4; n++) {
        const REAL *col = C + (size_t)(ix - 1 + n) * N1 + (iy - 1);
        s += wx[n] * (wy[0] * col[0] + wy[1] * col[1] + wy[2] * col[2] + wy[3] * col[3]);
    }
    return s;
}

/* dewarp!(out, dw, img) (src/dewarp.jl:233-248): the frame's cubic B-spline
 * interpolant (BSpline(Cubic(Line(OnGrid()))), extrapolated with zero)
 * sampled at the dewarper's precomputed source coordinates rows, cols
 * (ny x nx, 1-based, column-major, converted to the image precision T as
 * itp(T(rows), T(cols)) does); masked nodes (mask nullable) are 0. */
void F(orc_dewarp)(const REAL *img, int nr, int nc, const double *rows, const double *cols, const uint8_t *mask,
                   int ny, int nx, REAL *out)
{
    REAL *coef = (REAL *)malloc(sizeof(REAL) * (size_t)(nr + 2) * (nc + 2));
    F(orc_prefilter)(img, nr, nc, coef);
    for (size_t e = 0; e < (size_t)ny * nx; e++)
        out[e] = (mask && mask[e]) ? (REAL)0 : F(orc_cubic)(coef, nr, nc, (REAL)rows[e], (REAL)cols[e]);
    free(coef);
}

typedef struct {
    const REAL *cA, *cB, *pu, *pv;
    int nr, nc, gny, gnx;
    REAL y0, ysp, x0, xsp;
    REAL *wA, *wB;
} F(orc_deform_args);

static void F(orc_deform_range)(void *arg, int c0, int c1)
{
    F(orc_deform_args) *a = (F(orc_deform_args) *)arg;
    for (int c = c0; c < c1; c++)
        for (int r = 0; r < a->nr; r++) {
            REAL R = (REAL)(r + 1), C = (REAL)(c + 1);
            REAL du2 = F(orc_bilinear)(a->pu, a->gny, a->gnx, a->y0, a->ysp, a->x0, a->xsp, R, C) / 2;
            REAL dv2 = F(orc_bilinear)(a->pv, a->gny, a->gnx, a->y0, a->ysp, a->x0, a->xsp, R, C) / 2;
            a->wA[r + (size_t)c * a->nr] = F(orc_cubic)(a->cA, a->nr, a->nc, R - dv2, C - du2);
            a->wB[r + (size_t)c * a->nr] = F(orc_cubic)(a->cB, a->nr, a->nc, R + dv2, C + du2);
        }
}

/* deform_images / _ka_deform! (src/pipeline.jl:936-965,
 * src/ka_backend.jl:293-312). Knots py/px uniform. */
void F(orc_deform)(const REAL *cA, const REAL *cB, int nr, int nc, const REAL *py, int gny, const REAL *px, int gnx,
                   const REAL *pu, const REAL *pv, REAL *wA, REAL *wB, int nthreads)
{
    F(orc_deform_args) a;
    a.cA = cA; a.cB = cB; a.pu = pu; a.pv = pv;
    a.nr = nr; a.nc = nc; a.gny = gny; a.gnx = gnx;
    a.y0 = py[0]; a.x0 = px[0];
    a.ysp = gny > 1 ? py[1] - py[0] : (REAL)1;
    a.xsp = gnx > 1 ? px[1] - px[0] : (REAL)1;
    a.wA = wA; a.wB = wB;
    orc_parallel_for(nc, nthreads, F(orc_deform_range), &a);
}

/* predictor_node_values (src/pipeline.jl:561-569). */
void F(orc_node_values)(const REAL *py, int gny, const REAL *px, int gnx, const REAL *pu, const REAL *pv,
                        const REAL *y, int ny, const REAL *x, int nx, REAL *u, REAL *v)
{
    REAL y0 = py[0], x0 = px[0];
    REAL ysp = gny > 1 ? py[1] - py[0] : (REAL)1, xsp = gnx > 1 ? px[1] - px[0] : (REAL)1;
    for (int j = 0; j < nx; j++)
        for (int i = 0; i < ny; i++) {
            u[i + (size_t)j * ny] = F(orc_bilinear)(pu, gny, gnx, y0, ysp, x0, xsp, y[i], x[j]);
            v[i + (size_t)j * ny] = F(orc_bilinear)(pv, gny, gnx, y0, ysp, x0, xsp, y[i], x[j]);
        }
}

/* ======================================================================
 * Window processing (src/pipeline.jl:972-1027)
 * ==================================================================== */
typedef struct {
    const hh_pass *p;
    const F(orc_corr) * c;
    const REAL *A, *B;
    int nr;
    const int32_t *jobs;
    int ny, nx;
    REAL *u, *v, *pr, *cm, *alt_u, *alt_v, *unc_u, *unc_v;
    int32_t *locs_out;
    REAL *vals_out;
    int32_t *found_out;
    REAL *planes;
    const uint8_t *mask; /* image-sized pixel mask, nullable */
    REAL *tie;           /* per grid cell: near-tie margin of the primary peak, nullable */
} F(orc_pw_args);

/* Near-tie margin of a window's primary peak (test diagnostics): the
 * smallest relative gap (P1 - R(q)) / P1 between the primary value and any
 * 8-neighbour q of its location or the secondary peak value. A margin
 * below the float32 resolution of the plane means another computation of
 * the same window may take the neighbour (or the secondary) as the peak. */
static REAL F(orc_tie_margin)(const REAL *R, int m1, int m2, const REAL *vals, const int32_t *locs, int found)
{
    if (found < 1 || !(vals[0] > 0)) return (REAL)1;
    REAL P1 = vals[0], g = (REAL)1;
    int r = locs[0] - 1, c = locs[1] - 1;
    for (int dc = -1; dc <= 1; dc++)
        for (int dr = -1; dr <= 1; dr++) {
            if (!dr && !dc) continue;
            int rr = r + dr, cc = c + dc;
            if (rr < 0 || rr >= m1 || cc < 0 || cc >= m2) continue;
            REAL q = (P1 - R[rr + (size_t)cc * m1]) / P1;
            if (q < g) g = q;
        }
    if (found >= 2) {
        REAL q = (P1 - vals[1]) / P1;
        if (q < g) g = q;
    }
    return g;
}

static void F(orc_pw_range)(void *arg, int j0, int j1)
{
    F(orc_pw_args) *a = (F(orc_pw_args) *)arg;
    const hh_pass *p = a->p;
    F(orc_scratch) s;
    F(orc_scratch_init)(&s, a->c);
    int K = p->n_peaks > 2 ? p->n_peaks : 2;
    REAL vals[64];
    int32_t locs[128];
    REAL altu[64], altv[64];
    int wr = a->c->wr, wc = a->c->wc;
    int mw = wr > wc ? wr : wc;
    REAL *wA = (REAL *)malloc(sizeof(REAL) * wr * wc), *wB = (REAL *)malloc(sizeof(REAL) * wr * wc);
    REAL *dC = (REAL *)malloc(sizeof(REAL) * mw * mw), *dCs = (REAL *)malloc(sizeof(REAL) * mw * mw);
    size_t ngrid = (size_t)a->ny * a->nx;
    size_t np = (size_t)a->c->m1 * a->c->m2;
    for (int jb = j0; jb < j1; jb++) {
        int gi = a->jobs[4 * jb], gj = a->jobs[4 * jb + 1], rs = a->jobs[4 * jb + 2], cs = a->jobs[4 * jb + 3];
        size_t g = (size_t)(gi - 1) + (size_t)(gj - 1) * a->ny;
        F(orc_plane)(a->c, &s, a->A, a->B, a->nr, rs, cs, a->mask);
        if (a->planes) memcpy(a->planes + np * jb, s.R, sizeof(REAL) * np);
        REAL out[4];
        int found = F(orc_analyze_plane)(s.R, a->c->m1, a->c->m2, p, vals, locs, out, altu, altv);
        a->u[g] = out[0];
        a->v[g] = out[1];
        a->pr[g] = out[2];
        a->cm[g] = out[3];
        if (a->alt_u) {
            int mlim = found < p->n_peaks ? found : p->n_peaks;
            for (int m = 2; m <= mlim; m++) {
                a->alt_u[g + (size_t)(m - 2) * ngrid] = altu[m - 2];
                a->alt_v[g + (size_t)(m - 2) * ngrid] = altv[m - 2];
            }
        }
        if (a->found_out) a->found_out[jb] = found;
        if (a->tie) a->tie[g] = F(orc_tie_margin)(s.R, a->c->m1, a->c->m2, vals, locs, found);
        if (a->locs_out)
            for (int m = 0; m < K; m++) {
                a->locs_out[2 * (K * jb + m)] = m < found ? locs[2 * m] : 0;
                a->locs_out[2 * (K * jb + m) + 1] = m < found ? locs[2 * m + 1] : 0;
                a->vals_out[K * jb + m] = m < found ? vals[m] : (REAL)NAN;
            }
        if (a->unc_u) {
            double st[2 * HH_UQ_NSTATS];
            memset(st, 0, sizeof st);
            F(orc_uq_accumulate)(st, a->A, a->B, a->nr, rs, cs, wr, wc, a->c->apod, wA, wB, dC, dCs, a->mask);
            a->unc_u[g] = F(orc_uq_finalize)(st);
            a->unc_v[g] = F(orc_uq_finalize)(st + HH_UQ_NSTATS);
        }
    }
    free(wA); free(wB); free(dC); free(dCs);
    F(orc_scratch_free)(&s);
}

/* process_windows! over a job list (jobs: int32 4 x njobs, 1-based).
 * Debug outputs (locs 2 x K x njobs, vals K x njobs, found, planes
 * m1 x m2 x njobs) and the pixel mask nullable. */
void F(orc_process_windows)(const hh_pass *p, const REAL *A, const REAL *B, int nr, int nc, const int32_t *jobs,
                            int njobs, int ny, int nx, REAL *u, REAL *v, REAL *pr, REAL *cm, REAL *alt_u,
                            REAL *alt_v, REAL *unc_u, REAL *unc_v, int32_t *locs, REAL *vals, int32_t *found,
                            REAL *planes, const uint8_t *mask, int nthreads, REAL *tie)
{
    (void)nc;
    F(orc_corr) c;
    F(orc_corr_init)(&c, p);
    F(orc_pw_args) a = {p, &c, A, B, nr, jobs, ny, nx, u, v, pr, cm, alt_u, alt_v, unc_u, unc_v,
                        locs, vals, found, planes, mask, tie};
    orc_parallel_for(njobs, nthreads, F(orc_pw_range), &a);
    F(orc_corr_free)(&c);
}

/* correlation_plane! of one window pair (subA: wr x wc, subB: the search
 * area, sr x sc). */
void F(orc_correlation_plane)(const hh_pass *p, const REAL *subA, const REAL *subB, REAL *R)
{
    F(orc_corr) c;
    F(orc_corr_init)(&c, p);
    F(orc_scratch) s;
    F(orc_scratch_init)(&s, &c);
    F(orc_plane_at)(&c, &s, subA, c.wr, 1, 1, subB, c.sr, 1, 1, NULL);
    memcpy(R, s.R, sizeof(REAL) * c.m1 * c.m2);
    F(orc_scratch_free)(&s);
    F(orc_corr_free)(&c);
}

/* The gain plane of a pass (empty when not padded): m1 x m2. */
void F(orc_gain)(const hh_pass *p, REAL *out)
{
    F(orc_corr) c;
    F(orc_corr_init)(&c, p);
    if (c.gain) memcpy(out, c.gain, sizeof(REAL) * c.m1 * c.m2);
    F(orc_corr_free)(&c);
}

typedef struct {
    const hh_pass *p;
    const REAL *A, *B, *apod;
    int nr, ny;
    const int32_t *jobs;
    REAL *unc_u, *unc_v;
    double *stats;
    const uint8_t *mask;
} F(orc_uq_args);

static void F(orc_uq_range)(void *arg, int j0, int j1)
{
    F(orc_uq_args) *a = (F(orc_uq_args) *)arg;
    int wr = a->p->window_size[0], wc = a->p->window_size[1];
    int mw = wr > wc ? wr : wc;
    REAL *wA = (REAL *)malloc(sizeof(REAL) * wr * wc), *wB = (REAL *)malloc(sizeof(REAL) * wr * wc);
    REAL *dC = (REAL *)malloc(sizeof(REAL) * mw * mw), *dCs = (REAL *)malloc(sizeof(REAL) * mw * mw);
    for (int jb = j0; jb < j1; jb++) {
        int gi = a->jobs[4 * jb], gj = a->jobs[4 * jb + 1], rs = a->jobs[4 * jb + 2], cs = a->jobs[4 * jb + 3];
        double st[2 * HH_UQ_NSTATS];
        memset(st, 0, sizeof st);
        F(orc_uq_accumulate)(st, a->A, a->B, a->nr, rs, cs, wr, wc, a->apod, wA, wB, dC, dCs, a->mask);
        size_t g = (size_t)(gi - 1) + (size_t)(gj - 1) * a->ny;
        if (a->unc_u) a->unc_u[g] = F(orc_uq_finalize)(st);
        if (a->unc_v) a->unc_v[g] = F(orc_uq_finalize)(st + HH_UQ_NSTATS);
        if (a->stats) memcpy(a->stats + (size_t)2 * HH_UQ_NSTATS * jb, st, sizeof st);
    }
    free(wA); free(wB); free(dC); free(dCs);
}

/* uncertainty_sweep! (src/pipeline.jl:873-899). stats nullable
 * (2 x 44 per job: component-major); mask nullable (image-sized). */
void F(orc_uncertainty_sweep)(const hh_pass *p, const REAL *A, const REAL *B, int nr, const int32_t *jobs, int njobs,
                              int ny, REAL *unc_u, REAL *unc_v, double *stats, const uint8_t *mask, int nthreads)
{
    REAL *apod = (REAL *)malloc(sizeof(REAL) * p->window_size[0] * p->window_size[1]);
    F(orc_apodization)(p->window_size[0], p->window_size[1], p->apodization, apod);
    F(orc_uq_args) a = {p, A, B, apod, nr, ny, jobs, unc_u, unc_v, stats, mask};
    orc_parallel_for(njobs, nthreads, F(orc_uq_range), &a);
    free(apod);
}

/* ======================================================================
 * Pass orchestration (src/pipeline.jl:361-449,529-555,655-866)
 * ==================================================================== */

/* pass_grid (src/pipeline.jl:529-555), unmasked: x (nx), y (ny), jobs
 * (4 x ny*nx, column-major over the grid). Returns 0 or -1 if the search
 * area exceeds the image. */
int F(orc_pass_grid)(const hh_pass *p, int nr, int nc, int *ny, int *nx, REAL *x, REAL *y, int32_t *jobs)
{
    int wr = p->window_size[0], wc = p->window_size[1];
    int sr = p->search_area_size[0], sc = p->search_area_size[1];
    int mr = (sr - wr) / 2, mc = (sc - wc) / 2;
    if (sr > nr || sc > nc) return -1;
    int r0 = 1 + mr, rstep = wr - p->overlap[0], rstop = nr - wr - mr + 1;
    int c0 = 1 + mc, cstep = wc - p->overlap[1], cstop = nc - wc - mc + 1;
    int nyy = orc_range_count(r0, rstep, rstop), nxx = orc_range_count(c0, cstep, cstop);
    *ny = nyy;
    *nx = nxx;
    if (x)
        for (int j = 0; j < nxx; j++) x[j] = (REAL)(c0 + j * cstep) + (REAL)((wc - 1) / 2.0);
    if (y)
        for (int i = 0; i < nyy; i++) y[i] = (REAL)(r0 + i * rstep) + (REAL)((wr - 1) / 2.0);
    if (jobs) {
        int k = 0;
        for (int j = 0; j < nxx; j++)
            for (int i = 0; i < nyy; i++) {
                jobs[4 * k] = i + 1;
                jobs[4 * k + 1] = j + 1;
                jobs[4 * k + 2] = r0 + i * rstep;
                jobs[4 * k + 3] = c0 + j * cstep;
                k++;
            }
    }
    return 0;
}

/* The grid-level mask of pass_grid (src/pipeline.jl:542-549): a window whose
 * masked-pixel fraction (of the window or of its search area) reaches
 * thr produces no vector. gmask: ny x nx, column-major. */
void F(orc_grid_mask)(const hh_pass *p, int nr, int nc, const uint8_t *mask, double thr, uint8_t *gmask)
{
    int ny, nx;
    if (F(orc_pass_grid)(p, nr, nc, &ny, &nx, NULL, NULL, NULL) != 0) return;
    int wr = p->window_size[0], wc = p->window_size[1];
    int sr = p->search_area_size[0], sc = p->search_area_size[1];
    int mr = (sr - wr) / 2, mc = (sc - wc) / 2;
    int r0 = 1 + mr, rstep = wr - p->overlap[0], c0 = 1 + mc, cstep = wc - p->overlap[1];
    for (int gj = 0; gj < nx; gj++)
        for (int gi = 0; gi < ny; gi++) {
            const int rs = r0 + gi * rstep, cs = c0 + gj * cstep;
            long nw = 0, ns = 0;
            for (int j = 0; j < wc; j++)
                for (int i = 0; i < wr; i++) nw += mask[(size_t)(rs - 1 + i) + (size_t)(cs - 1 + j) * nr];
            const int srs = rs - mr, scs = cs - mc;
            for (int j = 0; j < sc; j++)
                for (int i = 0; i < sr; i++) ns += mask[(size_t)(srs - 1 + i) + (size_t)(scs - 1 + j) * nr];
            const double fw = (double)nw / (double)(wr * wc), fs = (double)ns / (double)(sr * sc);
            gmask[gi + (size_t)gj * ny] = fw >= thr || fs >= thr;
        }
}

/* field_change (src/pipeline.jl:849-866): q95 of max(|du|, |dv|) over
 * unmasked cells, NaN<->finite flips count as Inf. */
REAL F(orc_field_change)(const REAL *u, const REAL *v, const REAL *pu, const REAL *pv, const uint8_t *mask, int n)
{
    REAL *buf = (REAL *)malloc(sizeof(REAL) * (n > 0 ? n : 1));
    int m = 0;
    for (int i = 0; i < n; i++) {
        if (mask && mask[i]) continue;
        REAL du = (REAL)fabs((double)(u[i] - pu[i])), dv = (REAL)fabs((double)(v[i] - pv[i]));
        if (isnan(du) || isnan(dv)) {
            if (!!isnan(u[i]) == !!isnan(pu[i]) && !!isnan(v[i]) == !!isnan(pv[i])) continue;
            buf[m++] = (REAL)INFINITY;
            continue;
        }
        buf[m++] = du > dv ? du : dv;
    }
    REAL r = m == 0 ? (REAL)0 : F(orc_quantile)(buf, m, 0.95);
    free(buf);
    return r;
}

typedef struct {
    int ny, nx;
    REAL *x, *y, *u, *v; /* predictor grid (smoothed) */
} F(orc_pred);

/* build_predictor (src/pipeline.jl:454-464): masked cells (NaN) are filled
 * from valid neighbours (replace_vectors!) on copies, then smoothed. */
static void F(orc_build_predictor)(const REAL *u, const REAL *v, const uint8_t *gmask, int ny, int nx, int smoothing,
                                   REAL *pu, REAL *pv)
{
    size_t ng = (size_t)ny * nx;
    int anyg = 0;
    if (gmask)
        for (size_t q = 0; q < ng; q++) anyg |= gmask[q];
    REAL *tu = (REAL *)malloc(sizeof(REAL) * ng), *tv = (REAL *)malloc(sizeof(REAL) * ng);
    memcpy(tu, u, sizeof(REAL) * ng);
    memcpy(tv, v, sizeof(REAL) * ng);
    if (anyg) F(orc_replace)(tu, tv, gmask, ny, nx);
    if (smoothing) {
        F(orc_smooth)(tu, ny, nx, pu);
        F(orc_smooth)(tv, ny, nx, pv);
    } else {
        memcpy(pu, tu, sizeof(REAL) * ng);
        memcpy(pv, tv, sizeof(REAL) * ng);
    }
    free(tu);
    free(tv);
}

/* run_piv(imgA, imgB, passes; mask, mask_threshold) (src/pipeline.jl:
 * 361-449) with piv_pass (src/pipeline.jl:655-832). mask: image-sized pixel
 * mask, nullable; omask (nullable) receives the final grid mask. Outputs the
 * final pass's PIVResult into caller arrays sized by the final grid (any may
 * be NULL except the ones needed). ouqpert (nullable, test diagnostics):
 * the final uncertainty again (sigma_u grid, then sigma_v grid), computed on
 * the final pass's images re-deformed with every predictor node moved by up
 * to 1e-5 px (when the pass deforms), rounded to float32 and scaled by
 * 1 + 2^-21 h per pixel (h in [-1, 1], fixed hashes of the node and pixel
 * indices): how far sigma moves when its inputs carry the error of a
 * float32 pipeline. ouqstats (nullable, test diagnostics): the FP64 Wieneke
 * statistics behind the final uncertainty, 2 x 44 per grid cell (u then v;
 * zero at cells without a job). oties (nullable, test diagnostics): the
 * windows of every pass and sweep whose primary peak is a near-tie (margin
 * orc_tie_margin below tie_rel), as (pass, sweep, y, x) rows, at most
 * maxties; *nties receives their count.
 * Returns 0, or -1 for an invalid schedule. */
int F(orc_run_piv)(const REAL *imgA, const REAL *imgB, int nr, int nc, const hh_pass *passes, int npasses,
                   int predictor_smoothing, int nthreads, REAL *ox, REAL *oy, REAL *ou, REAL *ov, REAL *opr,
                   REAL *ocm, REAL *ouu, REAL *ouv, uint8_t *oout, const uint8_t *mask, double mask_threshold,
                   uint8_t *omask, REAL *ouqpert, double *ouqstats, double tie_rel, double *oties, int maxties,
                   int *nties)
{
    if (nties) *nties = 0;
    if (!nties) oties = NULL;
    if (npasses < 1) return -1;
    int deforms = npasses > 1;
    for (int k = 0; k < npasses; k++) deforms |= passes[k].max_iterations > 1;
    size_t npx = (size_t)nr * nc;
    REAL *cA = NULL, *cB = NULL, *wA = NULL, *wB = NULL;
    if (deforms) {
        cA = (REAL *)malloc(sizeof(REAL) * (size_t)(nr + 2) * (nc + 2));
        cB = (REAL *)malloc(sizeof(REAL) * (size_t)(nr + 2) * (nc + 2));
        wA = (REAL *)malloc(sizeof(REAL) * npx);
        wB = (REAL *)malloc(sizeof(REAL) * npx);
        F(orc_prefilter)(imgA, nr, nc, cA);
        F(orc_prefilter)(imgB, nr, nc, cB);
    }
    F(orc_pred) pred = {0, 0, NULL, NULL, NULL, NULL};
    int have_pred = 0;
    /* previous result */
    int pny = 0, pnx = 0;
    REAL *px = NULL, *py = NULL, *pu = NULL, *pv = NULL;
    uint8_t *pmask = NULL; /* the previous pass's grid mask */
    int rc = 0;
    for (int k = 0; k < npasses; k++) {
        const hh_pass *p = &passes[k];
        int force_replace = k < npasses - 1;
        if (px) {
            /* build_predictor (src/pipeline.jl:454-464), unmasked */
            pred.ny = pny;
            pred.nx = pnx;
            free(pred.x); free(pred.y); free(pred.u); free(pred.v);
            pred.x = px; pred.y = py;
            pred.u = (REAL *)malloc(sizeof(REAL) * pny * pnx);
            pred.v = (REAL *)malloc(sizeof(REAL) * pny * pnx);
            F(orc_build_predictor)(pu, pv, pmask, pny, pnx, predictor_smoothing, pred.u, pred.v);
            free(pu); free(pv); free(pmask);
            px = py = pu = pv = NULL;
            pmask = NULL;
            have_pred = 1;
        }
        int ny, nx;
        if (F(orc_pass_grid)(p, nr, nc, &ny, &nx, NULL, NULL, NULL) != 0) {
            rc = -1;
            break;
        }
        size_t ng = (size_t)ny * nx;
        REAL *x = (REAL *)malloc(sizeof(REAL) * nx), *y = (REAL *)malloc(sizeof(REAL) * ny);
        int32_t *jobs = (int32_t *)malloc(sizeof(int32_t) * 4 * (ng > 0 ? ng : 1));
        F(orc_pass_grid)(p, nr, nc, &ny, &nx, x, y, jobs);
        /* grid mask: masked windows produce no job */
        uint8_t *gmask = (uint8_t *)calloc(ng > 0 ? ng : 1, 1);
        int anyg = 0, njobs = (int)ng;
        if (mask) {
            F(orc_grid_mask)(p, nr, nc, mask, mask_threshold, gmask);
            njobs = 0;
            for (size_t q = 0; q < ng; q++) {
                anyg |= gmask[q];
                if (gmask[q]) continue;
                memmove(jobs + 4 * (size_t)njobs, jobs + 4 * q, 4 * sizeof(int32_t));
                njobs++;
            }
        }
        int maxiter = p->max_iterations;
        REAL *ru = (REAL *)calloc(ng, sizeof(REAL)), *rv = (REAL *)calloc(ng, sizeof(REAL));
        REAL *pr = (REAL *)calloc(ng, sizeof(REAL)), *cm = (REAL *)calloc(ng, sizeof(REAL));
        int nalt = p->n_peaks - 1;
        REAL *au = nalt > 0 ? (REAL *)malloc(sizeof(REAL) * ng * nalt) : NULL;
        REAL *av = nalt > 0 ? (REAL *)malloc(sizeof(REAL) * ng * nalt) : NULL;
        int unc = p->uncertainty && !force_replace;
        int fused_unc = unc && maxiter == 1;
        REAL *uu = (REAL *)malloc(sizeof(REAL) * ng), *uv = (REAL *)malloc(sizeof(REAL) * ng);
        for (size_t q = 0; q < ng; q++) uu[q] = uv[q] = (REAL)NAN;
        int keep_measured = maxiter > 1 && !(force_replace || p->replace_outliers);
        REAL *mu_ = keep_measured ? (REAL *)malloc(sizeof(REAL) * ng) : NULL;
        REAL *mv_ = keep_measured ? (REAL *)malloc(sizeof(REAL) * ng) : NULL;
        REAL *prevu = NULL, *prevv = NULL;
        REAL *u = (REAL *)malloc(sizeof(REAL) * ng), *v = (REAL *)malloc(sizeof(REAL) * ng);
        uint8_t *outl = (uint8_t *)calloc(ng > 0 ? ng : 1, 1);
        const REAL *srcA = imgA, *srcB = imgB;
        for (int it = 1; it <= maxiter; it++) {
            if (have_pred) {
                F(orc_deform)(cA, cB, nr, nc, pred.y, pred.ny, pred.x, pred.nx, pred.u, pred.v, wA, wB, nthreads);
                F(orc_node_values)(pred.y, pred.ny, pred.x, pred.nx, pred.u, pred.v, y, ny, x, nx, u, v);
                srcA = wA;
                srcB = wB;
            } else {
                memset(u, 0, sizeof(REAL) * ng);
                memset(v, 0, sizeof(REAL) * ng);
                srcA = imgA;
                srcB = imgB;
            }
            memset(ru, 0, sizeof(REAL) * ng);
            memset(rv, 0, sizeof(REAL) * ng);
            if (au)
                for (size_t q = 0; q < ng * nalt; q++) au[q] = av[q] = (REAL)NAN;
            REAL *tie = oties ? (REAL *)malloc(sizeof(REAL) * (ng > 0 ? ng : 1)) : NULL;
            if (tie)
                for (size_t q = 0; q < ng; q++) tie[q] = (REAL)1;
            F(orc_process_windows)(p, srcA, srcB, nr, nc, jobs, njobs, ny, nx, ru, rv, pr, cm, au, av,
                                   fused_unc ? uu : NULL, fused_unc ? uv : NULL, NULL, NULL, NULL, NULL, mask,
                                   nthreads, tie);
            if (tie) {
                for (size_t q = 0; q < ng; q++)
                    if ((double)tie[q] < tie_rel) {
                        if (*nties < maxties) {
                            double *row = oties + 4 * (size_t)*nties;
                            row[0] = k;
                            row[1] = it;
                            row[2] = (double)y[q % ny];
                            row[3] = (double)x[q / ny];
                        }
                        (*nties)++;
                    }
                free(tie);
            }
            if (au)
                for (int m = 0; m < nalt; m++)
                    for (size_t q = 0; q < ng; q++) {
                        au[q + m * ng] += u[q];
                        av[q + m * ng] += v[q];
                    }
            for (size_t q = 0; q < ng; q++) {
                u[q] += ru[q];
                v[q] += rv[q];
            }
            if (anyg)
                for (size_t q = 0; q < ng; q++)
                    if (gmask[q]) u[q] = v[q] = pr[q] = cm[q] = (REAL)NAN;
            if (keep_measured) {
                memcpy(mu_, u, sizeof(REAL) * ng);
                memcpy(mv_, v, sizeof(REAL) * ng);
            }
            F(orc_validate_and_replace)(p, ny, nx, force_replace || maxiter > 1, u, v, pr, cm, au, av,
                                        anyg ? gmask : NULL, outl);
            if (it == maxiter) break;
            if (p->convergence_tol > 0) {
                if (!prevu) {
                    prevu = (REAL *)malloc(sizeof(REAL) * ng);
                    prevv = (REAL *)malloc(sizeof(REAL) * ng);
                    memcpy(prevu, u, sizeof(REAL) * ng);
                    memcpy(prevv, v, sizeof(REAL) * ng);
                } else {
                    REAL ch = F(orc_field_change)(u, v, prevu, prevv, anyg ? gmask : NULL, (int)ng);
                    if (ch < (REAL)p->convergence_tol) break;
                    memcpy(prevu, u, sizeof(REAL) * ng);
                    memcpy(prevv, v, sizeof(REAL) * ng);
                }
            }
            /* build_predictor from this sweep's result */
            if (!have_pred) {
                pred.x = (REAL *)malloc(sizeof(REAL) * nx);
                pred.y = (REAL *)malloc(sizeof(REAL) * ny);
                pred.u = (REAL *)malloc(sizeof(REAL) * ng);
                pred.v = (REAL *)malloc(sizeof(REAL) * ng);
            } else if (pred.ny != ny || pred.nx != nx) {
                free(pred.x); free(pred.y); free(pred.u); free(pred.v);
                pred.x = (REAL *)malloc(sizeof(REAL) * nx);
                pred.y = (REAL *)malloc(sizeof(REAL) * ny);
                pred.u = (REAL *)malloc(sizeof(REAL) * ng);
                pred.v = (REAL *)malloc(sizeof(REAL) * ng);
            }
            pred.ny = ny;
            pred.nx = nx;
            memcpy(pred.x, x, sizeof(REAL) * nx);
            memcpy(pred.y, y, sizeof(REAL) * ny);
            F(orc_build_predictor)(u, v, anyg ? gmask : NULL, ny, nx, predictor_smoothing, pred.u, pred.v);
            have_pred = 1;
            if (!deforms) break; /* unreachable: iterating passes deform */
        }
        if (unc && !fused_unc)
            F(orc_uncertainty_sweep)(p, srcA, srcB, nr, jobs, njobs, ny, uu, uv, NULL, mask, nthreads);
        if (unc && ouqstats && k == npasses - 1) {
            double *st = (double *)calloc((size_t)2 * HH_UQ_NSTATS * (njobs > 0 ? njobs : 1), sizeof(double));
            F(orc_uncertainty_sweep)(p, srcA, srcB, nr, jobs, njobs, ny, NULL, NULL, st, mask, nthreads);
            memset(ouqstats, 0, sizeof(double) * 2 * HH_UQ_NSTATS * ng);
            for (int jb = 0; jb < njobs; jb++) {
                size_t g = (size_t)(jobs[4 * jb] - 1) + (size_t)(jobs[4 * jb + 1] - 1) * ny;
                memcpy(ouqstats + (size_t)2 * HH_UQ_NSTATS * g, st + (size_t)2 * HH_UQ_NSTATS * jb,
                       sizeof(double) * 2 * HH_UQ_NSTATS);
            }
            free(st);
        }
        if (unc && ouqpert && k == npasses - 1) {
            /* the final images again, from a predictor moved by up to 1e-5 px
             * per node (when the pass deforms), rounded to float32 and scaled
             * by 1 + 2^-21 h per pixel */
            REAL *qA = (REAL *)malloc(sizeof(REAL) * npx), *qB = (REAL *)malloc(sizeof(REAL) * npx);
            const REAL *sA = srcA, *sB = srcB;
            if (have_pred) {
                size_t nn = (size_t)pred.ny * pred.nx;
                REAL *tu = (REAL *)malloc(sizeof(REAL) * nn), *tv = (REAL *)malloc(sizeof(REAL) * nn);
                for (size_t q = 0; q < nn; q++) {
                    tu[q] = pred.u[q] + (REAL)(1e-5 * orc_hash_unit(2 * q + 7));
                    tv[q] = pred.v[q] + (REAL)(1e-5 * orc_hash_unit(2 * q + 8));
                }
                F(orc_deform)(cA, cB, nr, nc, pred.y, pred.ny, pred.x, pred.nx, tu, tv, qA, qB, nthreads);
                free(tu); free(tv);
                sA = qA;
                sB = qB;
            }
            for (size_t q = 0; q < npx; q++) {
                qA[q] = (REAL)((double)(float)sA[q] * (1.0 + ldexp(orc_hash_unit(2 * q), -21)));
                qB[q] = (REAL)((double)(float)sB[q] * (1.0 + ldexp(orc_hash_unit(2 * q + 1), -21)));
            }
            for (size_t q = 0; q < 2 * ng; q++) ouqpert[q] = (REAL)NAN;
            F(orc_uncertainty_sweep)(p, qA, qB, nr, jobs, njobs, ny, ouqpert, ouqpert + ng, NULL, mask, nthreads);
            free(qA); free(qB);
        }
        if (keep_measured)
            for (size_t q = 0; q < ng; q++)
                if (outl[q]) {
                    u[q] = mu_[q];
                    v[q] = mv_[q];
                }
        if (k == npasses - 1) {
            if (ox) memcpy(ox, x, sizeof(REAL) * nx);
            if (oy) memcpy(oy, y, sizeof(REAL) * ny);
            if (ou) memcpy(ou, u, sizeof(REAL) * ng);
            if (ov) memcpy(ov, v, sizeof(REAL) * ng);
            if (opr) memcpy(opr, pr, sizeof(REAL) * ng);
            if (ocm) memcpy(ocm, cm, sizeof(REAL) * ng);
            if (ouu) memcpy(ouu, uu, sizeof(REAL) * ng);
            if (ouv) memcpy(ouv, uv, sizeof(REAL) * ng);
            if (oout) memcpy(oout, outl, ng);
            if (omask) memcpy(omask, gmask, ng);
            free(x); free(y); free(u); free(v);
            free(gmask);
        } else {
            px = x; py = y; pu = u; pv = v;
            pny = ny; pnx = nx;
            pmask = gmask;
        }
        free(jobs); free(ru); free(rv); free(pr); free(cm); free(au); free(av);
        free(uu); free(uv); free(mu_); free(mv_); free(prevu); free(prevv); free(outl);
    }
    free(pred.x); free(pred.y); free(pred.u); free(pred.v);
    free(px); free(py); free(pu); free(pv); free(pmask);
    free(cA); free(cB); free(wA); free(wB);
    return rc;
}

/* ======================================================================
 * Ensemble (sum-of-correlation) PIV (src/ensemble.jl:47-343)
 * ==================================================================== */
typedef struct {
    const hh_pass *p;
    const F(orc_corr) * c;
    const REAL *A, *B;
    int nr;
    const int32_t *jobs;
    double *acc;  /* np per job */
    double *ust;  /* 2 x 44 per job, nullable */
    const uint8_t *mask; /* image-sized pixel mask, nullable */
} F(orc_ens_add_args);

/* accumulate_planes! (src/ensemble.jl:317-343): the pair's plane of each
 * window is added to that window's sum; with uncertainty, its Wieneke
 * statistics (window-sized B at the same origin) are pooled. Windows with
 * masked pixels load as in load_windows! (valid-pixel means, masked pixels
 * zero). */
static void F(orc_ens_add_range)(void *arg, int j0, int j1)
{
    F(orc_ens_add_args) *a = (F(orc_ens_add_args) *)arg;
    F(orc_scratch) s;
    F(orc_scratch_init)(&s, a->c);
    int wr = a->c->wr, wc = a->c->wc;
    int mw = wr > wc ? wr : wc;
    REAL *wA = (REAL *)malloc(sizeof(REAL) * wr * wc), *wB = (REAL *)malloc(sizeof(REAL) * wr * wc);
    REAL *dC = (REAL *)malloc(sizeof(REAL) * mw * mw), *dCs = (REAL *)malloc(sizeof(REAL) * mw * mw);
    size_t np = (size_t)a->c->m1 * a->c->m2;
    for (int jb = j0; jb < j1; jb++) {
        int rs = a->jobs[4 * jb + 2], cs = a->jobs[4 * jb + 3];
        F(orc_plane)(a->c, &s, a->A, a->B, a->nr, rs, cs, a->mask);
        double *dst = a->acc + np * jb;
        for (size_t e = 0; e < np; e++) dst[e] += (double)s.R[e];
        if (a->ust)
            F(orc_uq_accumulate)(a->ust + (size_t)2 * HH_UQ_NSTATS * jb, a->A, a->B, a->nr, rs, cs, wr, wc,
                                 a->c->apod, wA, wB, dC, dCs, a->mask);
    }
    free(wA); free(wB); free(dC); free(dCs);
    F(orc_scratch_free)(&s);
}

typedef struct {
    const hh_pass *p;
    const F(orc_corr) * c;
    const int32_t *jobs;
    int ny, nx;
    const double *acc, *ust;
    REAL *u, *v, *pr, *cm, *alt_u, *alt_v, *unc_u, *unc_v;
} F(orc_ens_an_args);

/* ensemble_analyze! (src/ensemble.jl:284-313): u, v hold the shared
 * predictor on entry and the total displacement on exit; alternatives are
 * predictor-relative; pooled statistics are finalized. */
static void F(orc_ens_an_range)(void *arg, int j0, int j1)
{
    F(orc_ens_an_args) *a = (F(orc_ens_an_args) *)arg;
    const hh_pass *p = a->p;
    size_t np = (size_t)a->c->m1 * a->c->m2;
    size_t ngrid = (size_t)a->ny * a->nx;
    REAL *R = (REAL *)malloc(sizeof(REAL) * np);
    REAL vals[64], altu[64], altv[64], out[4];
    int32_t locs[128];
    for (int jb = j0; jb < j1; jb++) {
        int gi = a->jobs[4 * jb], gj = a->jobs[4 * jb + 1];
        size_t g = (size_t)(gi - 1) + (size_t)(gj - 1) * a->ny;
        const double *src = a->acc + np * jb;
        for (size_t e = 0; e < np; e++) R[e] = (REAL)src[e];
        int found = F(orc_analyze_plane)(R, a->c->m1, a->c->m2, p, vals, locs, out, altu, altv);
        if (a->alt_u) {
            int mlim = found < p->n_peaks ? found : p->n_peaks;
            for (int m = 2; m <= mlim; m++) {
                a->alt_u[g + (size_t)(m - 2) * ngrid] = a->u[g] + altu[m - 2];
                a->alt_v[g + (size_t)(m - 2) * ngrid] = a->v[g] + altv[m - 2];
            }
        }
        a->u[g] += out[0];
        a->v[g] += out[1];
        a->pr[g] = out[2];
        a->cm[g] = out[3];
        if (a->ust) {
            a->unc_u[g] = F(orc_uq_finalize)(a->ust + (size_t)2 * HH_UQ_NSTATS * jb);
            a->unc_v[g] = F(orc_uq_finalize)(a->ust + (size_t)2 * HH_UQ_NSTATS * jb + HH_UQ_NSTATS);
        }
    }
    free(R);
}

/* run_piv_ensemble(pairs, passes; mask, mask_threshold) (src/ensemble.jl:
 * 47-84) with ensemble_pass (src/ensemble.jl:133-265). Each pass deforms
 * every pair by the shared predictor built from the previous pass, sums
 * each window's correlation planes over the pairs, analyzes the sums once
 * and validates (replacement forced on all but the last pass). The plane
 * sums are FP64 for either REAL (the reference sums in T); each sum is
 * rounded to REAL before its analysis. With a pixel mask (image-sized,
 * nullable) the grid mask removes windows as in run_piv, their cells are
 * NaN and validated as masked, and the next pass's predictor fills them
 * (src/ensemble.jl:154,254-264). planes_out (nullable) receives the last
 * pass's sums, m1 x m2 per job (the unmasked windows in column-major
 * order); omask (nullable) the final grid mask. Returns 0, or -1 for an
 * invalid schedule or an empty pair list. */
int F(orc_run_piv_ensemble)(const REAL *const *imgA, const REAL *const *imgB, int npairs, int nr, int nc,
                            const hh_pass *passes, int npasses, int predictor_smoothing, int nthreads, REAL *ox,
                            REAL *oy, REAL *ou, REAL *ov, REAL *opr, REAL *ocm, REAL *ouu, REAL *ouv, uint8_t *oout,
                            double *planes_out, const uint8_t *mask, double mask_threshold, uint8_t *omask)
{
    if (npasses < 1 || npairs < 1) return -1;
    size_t npx = (size_t)nr * nc;
    REAL *cA = NULL, *cB = NULL, *wA = NULL, *wB = NULL;
    if (npasses > 1) {
        cA = (REAL *)malloc(sizeof(REAL) * (size_t)(nr + 2) * (nc + 2));
        cB = (REAL *)malloc(sizeof(REAL) * (size_t)(nr + 2) * (nc + 2));
        wA = (REAL *)malloc(sizeof(REAL) * npx);
        wB = (REAL *)malloc(sizeof(REAL) * npx);
    }
    F(orc_pred) pred = {0, 0, NULL, NULL, NULL, NULL};
    int pny = 0, pnx = 0;
    REAL *px = NULL, *py = NULL, *pu = NULL, *pv = NULL;
    uint8_t *pmask = NULL; /* the previous pass's grid mask */
    int rc = 0;
    for (int k = 0; k < npasses; k++) {
        const hh_pass *p = &passes[k];
        int force_replace = k < npasses - 1;
        int have_pred = px != NULL;
        if (have_pred) {
            /* build_predictor (src/pipeline.jl:454-464) */
            free(pred.x); free(pred.y); free(pred.u); free(pred.v);
            pred.ny = pny;
            pred.nx = pnx;
            pred.x = px;
            pred.y = py;
            pred.u = (REAL *)malloc(sizeof(REAL) * pny * pnx);
            pred.v = (REAL *)malloc(sizeof(REAL) * pny * pnx);
            F(orc_build_predictor)(pu, pv, pmask, pny, pnx, predictor_smoothing, pred.u, pred.v);
            free(pu); free(pv); free(pmask);
            px = py = pu = pv = NULL;
            pmask = NULL;
        }
        int ny, nx;
        if (F(orc_pass_grid)(p, nr, nc, &ny, &nx, NULL, NULL, NULL) != 0) {
            rc = -1;
            break;
        }
        size_t ng = (size_t)ny * nx;
        REAL *x = (REAL *)malloc(sizeof(REAL) * nx), *y = (REAL *)malloc(sizeof(REAL) * ny);
        int32_t *jobs = (int32_t *)malloc(sizeof(int32_t) * 4 * (ng > 0 ? ng : 1));
        F(orc_pass_grid)(p, nr, nc, &ny, &nx, x, y, jobs);
        /* grid mask: masked windows produce no job (pass_grid, src/pipeline.jl:542-553) */
        uint8_t *gmask = (uint8_t *)calloc(ng > 0 ? ng : 1, 1);
        int anyg = 0, njobs = (int)ng;
        if (mask) {
            F(orc_grid_mask)(p, nr, nc, mask, mask_threshold, gmask);
            njobs = 0;
            for (size_t q = 0; q < ng; q++) {
                anyg |= gmask[q];
                if (gmask[q]) continue;
                memmove(jobs + 4 * (size_t)njobs, jobs + 4 * q, 4 * sizeof(int32_t));
                njobs++;
            }
        }
        F(orc_corr) c;
        F(orc_corr_init)(&c, p);
        size_t np = (size_t)c.m1 * c.m2;
        double *acc = (double *)calloc(ng * np > 0 ? ng * np : 1, sizeof(double));
        int unc = p->uncertainty && !force_replace;
        double *ust = unc ? (double *)calloc(ng * 2 * HH_UQ_NSTATS > 0 ? ng * 2 * HH_UQ_NSTATS : 1, sizeof(double))
                          : NULL;
        REAL *u = (REAL *)calloc(ng > 0 ? ng : 1, sizeof(REAL)), *v = (REAL *)calloc(ng > 0 ? ng : 1, sizeof(REAL));
        for (int q = 0; q < npairs; q++) {
            const REAL *srcA = imgA[q], *srcB = imgB[q];
            if (have_pred) {
                F(orc_prefilter)(imgA[q], nr, nc, cA);
                F(orc_prefilter)(imgB[q], nr, nc, cB);
                F(orc_deform)(cA, cB, nr, nc, pred.y, pred.ny, pred.x, pred.nx, pred.u, pred.v, wA, wB, nthreads);
                srcA = wA;
                srcB = wB;
            }
            F(orc_ens_add_args) aa = {p, &c, srcA, srcB, nr, jobs, acc, ust, mask};
            orc_parallel_for(njobs, nthreads, F(orc_ens_add_range), &aa);
        }
        /* u, v = the shared predictor at the nodes (zero without one) */
        if (have_pred) F(orc_node_values)(pred.y, pred.ny, pred.x, pred.nx, pred.u, pred.v, y, ny, x, nx, u, v);
        REAL *pr = (REAL *)calloc(ng > 0 ? ng : 1, sizeof(REAL)), *cm = (REAL *)calloc(ng > 0 ? ng : 1, sizeof(REAL));
        REAL *uu = (REAL *)malloc(sizeof(REAL) * (ng > 0 ? ng : 1)), *uv = (REAL *)malloc(sizeof(REAL) * (ng > 0 ? ng : 1));
        for (size_t q = 0; q < ng; q++) uu[q] = uv[q] = (REAL)NAN;
        int nalt = p->n_peaks - 1;
        REAL *au = nalt > 0 ? (REAL *)malloc(sizeof(REAL) * ng * nalt) : NULL;
        REAL *av = nalt > 0 ? (REAL *)malloc(sizeof(REAL) * ng * nalt) : NULL;
        if (au)
            for (size_t q = 0; q < ng * nalt; q++) au[q] = av[q] = (REAL)NAN;
        F(orc_ens_an_args) an = {p, &c, jobs, ny, nx, acc, ust, u, v, pr, cm, au, av, uu, uv};
        orc_parallel_for(njobs, nthreads, F(orc_ens_an_range), &an);
        if (anyg)
            for (size_t q = 0; q < ng; q++)
                if (gmask[q]) u[q] = v[q] = pr[q] = cm[q] = (REAL)NAN;
        uint8_t *outl = (uint8_t *)calloc(ng > 0 ? ng : 1, 1);
        F(orc_validate_and_replace)(p, ny, nx, force_replace, u, v, pr, cm, au, av, anyg ? gmask : NULL, outl);
        if (k == npasses - 1) {
            if (ox) memcpy(ox, x, sizeof(REAL) * nx);
            if (oy) memcpy(oy, y, sizeof(REAL) * ny);
            if (ou) memcpy(ou, u, sizeof(REAL) * ng);
            if (ov) memcpy(ov, v, sizeof(REAL) * ng);
            if (opr) memcpy(opr, pr, sizeof(REAL) * ng);
            if (ocm) memcpy(ocm, cm, sizeof(REAL) * ng);
            if (ouu) memcpy(ouu, uu, sizeof(REAL) * ng);
            if (ouv) memcpy(ouv, uv, sizeof(REAL) * ng);
            if (oout) memcpy(oout, outl, ng);
            if (planes_out) memcpy(planes_out, acc, sizeof(double) * (size_t)njobs * np);
            if (omask) memcpy(omask, gmask, ng);
            free(x); free(y); free(u); free(v);
            free(gmask);
        } else {
            px = x; py = y; pu = u; pv = v;
            pny = ny; pnx = nx;
            pmask = anyg ? gmask : NULL;
            if (!anyg) free(gmask);
        }
        free(jobs); free(acc); free(ust); free(pr); free(cm); free(uu); free(uv); free(au); free(av); free(outl);
        F(orc_corr_free)(&c);
    }
    free(pred.x); free(pred.y); free(pred.u); free(pred.v);
    free(px); free(py); free(pu); free(pv); free(pmask);
    free(cA); free(cB); free(wA); free(wB);
    return rc;
}

/* ======================================================================
 * Windowed correlation energy and its adjoint (build-defined: the reference
 * has no gradient path; its closest relative is correlate_deformable,
 * src/correlators.jl:620-638). For a node field (pu, pv) on the uniform
 * knots (py, px), pixel displacements are the bilinear Flat() values of
 * the field (as apply_predictor, src/pipeline.jl:575-600); the pair is
 * warped symmetrically, A^(r,c) = A(r - v/2, c - u/2) and
 * B^(r,c) = B(r + v/2, c + u/2), with bilinear image sampling (zero outside
 * [1, n]); over the windows of pass p (pass_grid, apodization p_x),
 *     E = - sum_w sum_{x in w} p_x (A^_x - mean_w A^) (B^_x - mean_w B^).
 * ==================================================================== */

/* Bilinear sample of I (nr x nc, 1-based coordinates) and its partial
 * derivatives; zero (and zero slope) outside [1, nr] x [1, nc]. */
static REAL F(orc_bil_sample)(const REAL *I, int nr, int nc, REAL y, REAL x, REAL *sy, REAL *sx)
{
    if (y < 1 || y > (REAL)nr || x < 1 || x > (REAL)nc) {
        *sy = *sx = 0;
        return 0;
    }
    int i0 = (int)floor((double)y), j0 = (int)floor((double)x);
    if (i0 > nr - 1) i0 = nr - 1;
    if (j0 > nc - 1) j0 = nc - 1;
    REAL fy = y - (REAL)i0, fx = x - (REAL)j0;
    const REAL *q = I + (size_t)(i0 - 1) + (size_t)(j0 - 1) * nr;
    REAL a = q[0], b = q[nr], c = q[1], d = q[nr + 1]; /* (i0,j0) (i0,j0+1) (i0+1,j0) (i0+1,j0+1) */
    *sy = (1 - fx) * (c - a) + fx * (d - b);
    *sx = (1 - fy) * (b - a) + fy * (d - c);
    return (1 - fy) * ((1 - fx) * a + fx * b) + fy * ((1 - fx) * c + fx * d);
}

/* Cubic B-spline sample of the padded coefficients C ((nr+2) x (nc+2), as
 * orc_cubic) and its partial derivatives; zero outside [1, nr] x [1, nc]. */
static REAL F(orc_cub_sample)(const REAL *Cf, int nr, int nc, REAL y, REAL x, REAL *sy, REAL *sx)
{
    if (y < 1 || y > (REAL)nr || x < 1 || x > (REAL)nc) {
        *sy = *sx = 0;
        return 0;
    }
    int iy = (int)floor((double)y), ix = (int)floor((double)x);
    if (iy > nr - 1) iy = nr - 1;
    if (ix > nc - 1) ix = nc - 1;
    REAL dy = y - (REAL)iy, dx = x - (REAL)ix;
    REAL wy[4], wx[4], gy[4], gx[4];
    REAL ey = 1 - dy, ex = 1 - dx;
    wy[0] = ey * ey * ey / 6; wy[1] = (4 - 6 * dy * dy + 3 * dy * dy * dy) / 6;
    wy[2] = (1 + 3 * dy + 3 * dy * dy - 3 * dy * dy * dy) / 6; wy[3] = dy * dy * dy / 6;
    wx[0] = ex * ex * ex / 6; wx[1] = (4 - 6 * dx * dx + 3 * dx * dx * dx) / 6;
    wx[2] = (1 + 3 * dx + 3 * dx * dx - 3 * dx * dx * dx) / 6; wx[3] = dx * dx * dx / 6;
    gy[0] = -ey * ey / 2; gy[1] = -2 * dy + (REAL)1.5 * dy * dy; gy[2] = (REAL)0.5 + dy - (REAL)1.5 * dy * dy;
    gy[3] = dy * dy / 2;
    gx[0] = -ex * ex / 2; gx[1] = -2 * dx + (REAL)1.5 * dx * dx; gx[2] = (REAL)0.5 + dx - (REAL)1.5 * dx * dx;
    gx[3] = dx * dx / 2;
    int N1 = nr + 2;
    REAL s = 0, s_y = 0, s_x = 0;
    for (int n = 0; n < 4; n++) {
        const REAL *col = Cf + (size_t)(ix - 1 + n) * N1 + (iy - 1);
        REAL cv = wy[0] * col[0] + wy[1] * col[1] + wy[2] * col[2] + wy[3] * col[3];
        REAL cd = gy[0] * col[0] + gy[1] * col[1] + gy[2] * col[2] + gy[3] * col[3];
        s += wx[n] * cv;
        s_y += wx[n] * cd;
        s_x += gx[n] * cv;
    }
    *sy = s_y;
    *sx = s_x;
    return s;
}

/* E of the node field; with gu/gv (gny x gnx) non-NULL also dE/dpu, dE/dpv.
 * interp: 0 bilinear image sampling, 1 the cubic B-spline interpolant of
 * the deformation path (image_interpolant, src/pipeline.jl:909-915).
 * Returns E, or NAN for an invalid pass. */
double F(orc_grad_energy)(const hh_pass *p, const REAL *A, const REAL *B, int nr, int nc, int interp, const REAL *py,
                          int gny, const REAL *px, int gnx, const REAL *pu, const REAL *pv, REAL *gu, REAL *gv)
{
    int ny, nx;
    if (F(orc_pass_grid)(p, nr, nc, &ny, &nx, NULL, NULL, NULL) != 0) return NAN;
    size_t ng = (size_t)ny * nx, npx = (size_t)nr * nc;
    int32_t *jobs = (int32_t *)malloc(sizeof(int32_t) * 4 * (ng > 0 ? ng : 1));
    F(orc_pass_grid)(p, nr, nc, &ny, &nx, NULL, NULL, jobs);
    int wr = p->window_size[0], wc = p->window_size[1];
    REAL *apod = (REAL *)malloc(sizeof(REAL) * wr * wc);
    F(orc_apodization)(wr, wc, p->apodization, apod);
    REAL y0 = py[0], ysp = gny > 1 ? py[1] - py[0] : 1, x0 = px[0], xsp = gnx > 1 ? px[1] - px[0] : 1;
    /* per-pixel warp and slopes */
    REAL *Ah = (REAL *)malloc(sizeof(REAL) * npx), *Bh = (REAL *)malloc(sizeof(REAL) * npx);
    REAL *Ax = (REAL *)malloc(sizeof(REAL) * npx), *Ay = (REAL *)malloc(sizeof(REAL) * npx);
    REAL *Bx = (REAL *)malloc(sizeof(REAL) * npx), *By = (REAL *)malloc(sizeof(REAL) * npx);
    REAL *cA = NULL, *cB = NULL;
    if (interp) {
        cA = (REAL *)malloc(sizeof(REAL) * (size_t)(nr + 2) * (nc + 2));
        cB = (REAL *)malloc(sizeof(REAL) * (size_t)(nr + 2) * (nc + 2));
        F(orc_prefilter)(A, nr, nc, cA);
        F(orc_prefilter)(B, nr, nc, cB);
    }
    for (int c = 1; c <= nc; c++)
        for (int r = 1; r <= nr; r++) {
            size_t e = (size_t)(r - 1) + (size_t)(c - 1) * nr;
            REAL u = F(orc_bilinear)(pu, gny, gnx, y0, ysp, x0, xsp, (REAL)r, (REAL)c);
            REAL v = F(orc_bilinear)(pv, gny, gnx, y0, ysp, x0, xsp, (REAL)r, (REAL)c);
            if (interp) {
                Ah[e] = F(orc_cub_sample)(cA, nr, nc, (REAL)r - v / 2, (REAL)c - u / 2, &Ay[e], &Ax[e]);
                Bh[e] = F(orc_cub_sample)(cB, nr, nc, (REAL)r + v / 2, (REAL)c + u / 2, &By[e], &Bx[e]);
            } else {
                Ah[e] = F(orc_bil_sample)(A, nr, nc, (REAL)r - v / 2, (REAL)c - u / 2, &Ay[e], &Ax[e]);
                Bh[e] = F(orc_bil_sample)(B, nr, nc, (REAL)r + v / 2, (REAL)c + u / 2, &By[e], &Bx[e]);
            }
        }
    free(cA);
    free(cB);
    double *GA = gu ? (double *)calloc(npx, sizeof(double)) : NULL;
    double *GB = gu ? (double *)calloc(npx, sizeof(double)) : NULL;
    double E = 0, N = (double)wr * wc;
    for (size_t k = 0; k < ng; k++) {
        int rs = jobs[4 * k + 2], cs = jobs[4 * k + 3];
        double sa = 0, sb = 0;
        for (int j = 0; j < wc; j++)
            for (int i = 0; i < wr; i++) {
                size_t e = (size_t)(rs - 1 + i) + (size_t)(cs - 1 + j) * nr;
                sa += Ah[e];
                sb += Bh[e];
            }
        double mA = sa / N, mB = sb / N, PA = 0, PB = 0, Ew = 0;
        for (int j = 0; j < wc; j++)
            for (int i = 0; i < wr; i++) {
                size_t e = (size_t)(rs - 1 + i) + (size_t)(cs - 1 + j) * nr;
                double pw = apod[i + j * wr], a_ = Ah[e] - mA, b_ = Bh[e] - mB;
                PA += pw * a_;
                PB += pw * b_;
                Ew -= pw * a_ * b_;
            }
        E += Ew;
        if (!gu) continue;
        for (int j = 0; j < wc; j++)
            for (int i = 0; i < wr; i++) {
                size_t e = (size_t)(rs - 1 + i) + (size_t)(cs - 1 + j) * nr;
                double pw = apod[i + j * wr];
                GA[e] += -pw * (Bh[e] - mB) + PB / N;
                GB[e] += -pw * (Ah[e] - mA) + PA / N;
            }
    }
    if (gu) {
        for (size_t q = 0; q < (size_t)gny * gnx; q++) gu[q] = gv[q] = 0;
        double *nu = (double *)calloc((size_t)gny * gnx, sizeof(double));
        double *nv = (double *)calloc((size_t)gny * gnx, sizeof(double));
        for (int c = 1; c <= nc; c++)
            for (int r = 1; r <= nr; r++) {
                size_t e = (size_t)(r - 1) + (size_t)(c - 1) * nr;
                double gx = -0.5 * GA[e] * Ax[e] + 0.5 * GB[e] * Bx[e];
                double gyv = -0.5 * GA[e] * Ay[e] + 0.5 * GB[e] * By[e];
                int iy, ix;
                REAL ly, lx;
                F(orc_lin_axis)(y0, ysp, gny, (REAL)r, &iy, &ly);
                F(orc_lin_axis)(x0, xsp, gnx, (REAL)c, &ix, &lx);
                int iy2 = gny == 1 ? iy : iy + 1, ix2 = gnx == 1 ? ix : ix + 1;
                double w00 = (1 - (double)ly) * (1 - (double)lx), w01 = (1 - (double)ly) * lx;
                double w10 = (double)ly * (1 - (double)lx), w11 = (double)ly * lx;
                nu[iy + ix * gny] += w00 * gx;
                nu[iy + ix2 * gny] += w01 * gx;
                nu[iy2 + ix * gny] += w10 * gx;
                nu[iy2 + ix2 * gny] += w11 * gx;
                nv[iy + ix * gny] += w00 * gyv;
                nv[iy + ix2 * gny] += w01 * gyv;
                nv[iy2 + ix * gny] += w10 * gyv;
                nv[iy2 + ix2 * gny] += w11 * gyv;
            }
        for (size_t q = 0; q < (size_t)gny * gnx; q++) {
            gu[q] = (REAL)nu[q];
            gv[q] = (REAL)nv[q];
        }
        free(nu); free(nv);
    }
    free(jobs); free(apod); free(Ah); free(Bh); free(Ax); free(Ay); free(Bx); free(By); free(GA); free(GB);
    return E;
}

#undef F
#undef REAL_EPS
#undef RLOG
