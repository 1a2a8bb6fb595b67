/*
 * FP64 CPU oracle for the direct J/K build -- TEST INFRASTRUCTURE ONLY.
 * See oracle.h for scope, conventions and the ChronusQ files restated.
 */
#include "oracle.h"

#include <complex.h>
#include <ctype.h>
#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MAXL 6          /* highest shell l handled */
#define MAXPRIM 32
#define NCART(l) (((l) + 1) * ((l) + 2) / 2)
#define NOFF(L) ((L) * ((L) + 1) * ((L) + 2) / 6) /* #cart with l < L */

static __thread char g_err[512];
const char *oracle_last_error(void) { return g_err; }
static void set_err(const char *m) { snprintf(g_err, sizeof g_err, "%s", m); }

/* ------------------------------------------------------------------ */
/* elements                                                            */
/* ------------------------------------------------------------------ */
static const char *ELEM[] = {
    "X",  "H",  "He", "Li", "Be", "B",  "C",  "N",  "O",  "F",  "Ne", "Na",
    "Mg", "Al", "Si", "P",  "S",  "Cl", "Ar", "K",  "Ca", "Sc", "Ti", "V",
    "Cr", "Mn", "Fe", "Co", "Ni", "Cu", "Zn", "Ga", "Ge", "As", "Se", "Br",
    "Kr", "Rb", "Sr", "Y",  "Zr", "Nb", "Mo", "Tc", "Ru", "Rh", "Pd", "Ag",
    "Cd", "In", "Sn", "Sb", "Te", "I",  "Xe"};
#define NELEM ((int)(sizeof ELEM / sizeof ELEM[0]))

static int elem_z(const char *s) {
  for (int z = 1; z < NELEM; z++) {
    const char *e = ELEM[z];
    size_t n = strlen(e);
    if (strlen(s) != n) continue;
    int ok = 1;
    for (size_t i = 0; i < n; i++)
      if (tolower((unsigned char)s[i]) != tolower((unsigned char)e[i])) ok = 0;
    if (ok) return z;
  }
  return -1;
}

/* ------------------------------------------------------------------ */
/* math helpers                                                        */
/* ------------------------------------------------------------------ */
static double dfact(int n) { /* n!!, (-1)!! = 1 */
  double r = 1.0;
  for (int k = n; k > 1; k -= 2) r *= k;
  return r;
}
static double binom(int n, int k) {
  if (k < 0 || k > n) return 0.0;
  double r = 1.0;
  for (int i = 1; i <= k; i++) r = r * (n - k + i) / i;
  return r;
}
/* position of (lx,ly,lz) inside its l block: lx descending, ly descending */
static int cart_pos(int lx, int ly, int lz) {
  int l = lx + ly + lz;
  return (l - lx) * (l - lx + 1) / 2 + lz;
}
typedef struct { int x, y, z; } ijk_t;
static ijk_t CART[NOFF(2 * MAXL + 2)]; /* combined index -> exponents */
static double SPH[MAXL + 1][2 * MAXL + 1][NCART(MAXL)];
static int g_tables_ready = 0;

static void build_tables(void) {
  if (g_tables_ready) return;
#pragma omp critical(oracle_tables)
  {
    if (!g_tables_ready) {
      for (int L = 0; L <= 2 * MAXL + 1; L++)
        for (int lx = L; lx >= 0; lx--)
          for (int ly = L - lx; ly >= 0; ly--) {
            int lz = L - lx - ly;
            ijk_t t = {lx, ly, lz};
            CART[NOFF(L) + cart_pos(lx, ly, lz)] = t;
          }
      /* Real solid harmonics in Cartesian monomials (Helgaker, Jorgensen,
       * Olsen, Molecular Electronic-Structure Theory, eq. 6.4.47-6.4.50),
       * re-normalised against x^l-normalised Cartesian Gaussians. */
      for (int l = 0; l <= MAXL; l++) {
        for (int m = -l; m <= l; m++) {
          double *c = SPH[l][m + l];
          for (int k = 0; k < NCART(l); k++) c[k] = 0.0;
          int am = abs(m);
          for (int t = 0; t <= (l - am) / 2; t++)
            for (int u = 0; u <= t; u++)
              for (int k2 = (m < 0 ? 1 : 0); k2 <= am; k2 += 2) {
                int sgn = ((t + (k2 - (m < 0 ? 1 : 0)) / 2) % 2) ? -1 : 1;
                double v = sgn * pow(0.25, t) * binom(l, t) *
                           binom(l - t, am + t) * binom(t, u) * binom(am, k2);
                int px = 2 * t + am - 2 * u - k2, py = 2 * u + k2,
                    pz = l - 2 * t - am;
                c[cart_pos(px, py, pz)] += v;
              }
          double nrm = 0.0;
          for (int a = 0; a < NCART(l); a++)
            for (int b = 0; b < NCART(l); b++) {
              ijk_t A = CART[NOFF(l) + a], B = CART[NOFF(l) + b];
              int sx = A.x + B.x, sy = A.y + B.y, sz = A.z + B.z;
              if ((sx | sy | sz) & 1) continue;
              nrm += c[a] * c[b] * dfact(sx - 1) * dfact(sy - 1) * dfact(sz - 1);
            }
          nrm /= dfact(2 * l - 1);
          double f = 1.0 / sqrt(nrm);
          for (int k = 0; k < NCART(l); k++) c[k] *= f;
        }
      }
      g_tables_ready = 1;
    }
  }
}

/* ------------------------------------------------------------------ */
/* Boys function                                                       */
/* ------------------------------------------------------------------ */
/* F_m(T) = int_0^1 t^{2m} exp(-T t^2) dt.  Below T = 36 the highest order is
 * summed from the all-positive series exp(-T) sum_k (2T)^k / (2m+1)(2m+3)..
 * (2m+2k+1) and lower orders follow by downward recursion; above, F_0 from
 * erf and upward recursion (stable once T > m + 1/2). */
void oracle_boys(int mmax, double T, double *F) {
  if (T < 36.0) {
    double et = exp(-T);
    double term = 1.0 / (2 * mmax + 1), sum = term;
    for (int k = 1; k < 2000; k++) {
      term *= 2.0 * T / (2 * mmax + 2 * k + 1);
      sum += term;
      if (term < 1e-18 * sum) break;
    }
    F[mmax] = et * sum;
    for (int m = mmax; m > 0; m--) F[m - 1] = (2.0 * T * F[m] + et) / (2 * m - 1);
  } else {
    double et = exp(-T);
    F[0] = 0.5 * sqrt(M_PI / T) * erf(sqrt(T));
    for (int m = 0; m < mmax; m++) F[m + 1] = ((2 * m + 1) * F[m] - et) / (2.0 * T);
  }
}

/* ------------------------------------------------------------------ */
/* basis                                                               */
/* ------------------------------------------------------------------ */
typedef struct {
  int l, pure, nprim, atom, bf, nfunc;
  double O[3];
  double alpha[MAXPRIM], coef[MAXPRIM];
} oshell;

typedef struct { int l, pure, nprim; double alpha[MAXPRIM], coef[MAXPRIM]; } rawshell;
typedef struct { int z; int n; rawshell *sh; } elemrec;

struct oracle_basis {
  int nshell, nbf, maxprim, maxl, natom;
  oshell *sh;
  int npairs;
  int *sig_off;   /* nshell+1 offsets into sig_s2 */
  int *sig_s2;    /* ascending s2 <= s1 */
  double *schwarz;
  long long st_computed, st_skipped;
};

static double str2d(const char *s) {
  char buf[64];
  size_t n = strlen(s);
  if (n >= sizeof buf) n = sizeof buf - 1;
  for (size_t i = 0; i < n; i++) buf[i] = (s[i] == 'D' || s[i] == 'd') ? 'E' : s[i];
  buf[n] = 0;
  return strtod(buf, NULL);
}

static int lmap(const char *s) {
  static const char *sym = "SPDFGHIJKLMNOQRTUVWXYZ";
  if (strlen(s) != 1) return -1;
  const char *p = strchr(sym, toupper((unsigned char)s[0]));
  return p ? (int)(p - sym) : -1;
}

/* next line from text; returns 0 at end.  Line is right-trimmed. */
static int next_line(const char **cur, char *line, size_t cap) {
  const char *p = *cur;
  if (!*p) return 0;
  size_t n = 0;
  while (*p && *p != '\n') {
    if (n + 1 < cap) line[n++] = *p;
    p++;
  }
  if (*p == '\n') p++;
  line[n] = 0;
  while (n > 0 && isspace((unsigned char)line[n - 1])) line[--n] = 0;
  *cur = p;
  return 1;
}
static int split(char *line, char **tok, int maxt) {
  int n = 0;
  char *s = strtok(line, " \t\r");
  while (s && n < maxt) {
    tok[n++] = s;
    s = strtok(NULL, " \t\r");
  }
  return n;
}

/* G94 parsing rules of ReferenceBasisSet::parseBasisFile
 * (src/basisset/reference.cxx:127-289): "****" opens a record whose first
 * line names the element; SP shells become an S and a P shell sharing
 * exponents; pure functions iff l > 1 unless Cartesian is forced. */
static int parse_g94(const char *text, int force_cart, elemrec *recs, int maxrec) {
  const char *cur = text;
  char line[512], lcopy[512];
  int nrec = 0, inrec = 0, newrec = 0;
  elemrec *r = NULL;
  while (next_line(&cur, line, sizeof line)) {
    const char *t = line;
    while (isspace((unsigned char)*t)) t++;
    if (!*t) continue;
    if (*t == '!') continue;
    if (!strcmp(t, "****")) {
      if (!next_line(&cur, line, sizeof line)) { inrec = 0; break; }
      t = line;
      while (isspace((unsigned char)*t)) t++;
      if (!*t) { inrec = 0; continue; }
      inrec = 1;
      newrec = 1;
    }
    if (!inrec) continue;
    strcpy(lcopy, t);
    char *tok[8];
    int nt = split(lcopy, tok, 8);
    if (nt == 0) continue;
    if (newrec) {
      if (nrec >= maxrec) return -1;
      r = &recs[nrec++];
      r->z = elem_z(tok[0]);
      r->n = 0;
      r->sh = NULL;
      newrec = 0;
      continue;
    }
    if (nt < 2) return -1;
    int sp = !strcmp(tok[0], "SP") || !strcmp(tok[0], "sp");
    int depth = atoi(tok[1]);
    int L = sp ? 0 : lmap(tok[0]);
    if (L < 0 || L > MAXL || depth <= 0 || depth > MAXPRIM) return -1;
    rawshell a, b;
    memset(&a, 0, sizeof a);
    memset(&b, 0, sizeof b);
    for (int i = 0; i < depth; i++) {
      if (!next_line(&cur, line, sizeof line)) return -1;
      strcpy(lcopy, line);
      char *pt[4];
      int np = split(lcopy, pt, 4);
      if (np < (sp ? 3 : 2)) return -1;
      a.alpha[i] = b.alpha[i] = str2d(pt[0]);
      a.coef[i] = str2d(pt[1]);
      if (sp) b.coef[i] = str2d(pt[2]);
    }
    a.nprim = b.nprim = depth;
    a.l = L;
    a.pure = (L > 1) && !force_cart;
    b.l = 1;
    b.pure = 0;
    int add = sp ? 2 : 1;
    r->sh = realloc(r->sh, sizeof(rawshell) * (r->n + add));
    r->sh[r->n++] = a;
    if (sp) r->sh[r->n++] = b;
  }
  return nrec;
}

/* libint2-convention normalisation: each primitive normalised as x^l, then
 * the contraction rescaled to unit self-overlap. */
static void normalize_shell(oshell *s) {
  const double pi32 = pow(M_PI, 1.5);
  int l = s->l;
  for (int p = 0; p < s->nprim; p++) {
    double ta = 2.0 * s->alpha[p];
    s->coef[p] *= sqrt(pow(2.0, l) * pow(ta, l + 1.5) / (pi32 * dfact(2 * l - 1)));
  }
  double nrm = 0.0;
  for (int p = 0; p < s->nprim; p++)
    for (int q = 0; q < s->nprim; q++) {
      double g = s->alpha[p] + s->alpha[q];
      nrm += dfact(2 * l - 1) * pi32 * s->coef[p] * s->coef[q] /
             (pow(2.0, l) * pow(g, l + 1.5));
    }
  double f = 1.0 / sqrt(nrm);
  for (int p = 0; p < s->nprim; p++) s->coef[p] *= f;
}

static void compute_pairs(oracle_basis *b);

oracle_basis *oracle_basis_create(const char *gbs, int natom, const int *Z,
                                  const double *xyz, int force_cart) {
  build_tables();
  elemrec recs[128];
  int nrec = parse_g94(gbs, force_cart, recs, 128);
  if (nrec < 0) { set_err("G94 parse error"); return NULL; }
  oracle_basis *b = calloc(1, sizeof *b);
  b->natom = natom;
  int cap = 0;
  for (int a = 0; a < natom; a++) {
    elemrec *r = NULL;
    for (int i = 0; i < nrec; i++)  /* later records override earlier */
      if (recs[i].z == Z[a]) r = &recs[i];
    if (!r) {
      snprintf(g_err, sizeof g_err, "Cannot find Z=%d in Basis Definition", Z[a]);
      free(b->sh);
      free(b);
      for (int i = 0; i < nrec; i++) free(recs[i].sh);
      return NULL;
    }
    for (int k = 0; k < r->n; k++) {
      if (b->nshell == cap) {
        cap = cap ? 2 * cap : 64;
        b->sh = realloc(b->sh, sizeof(oshell) * cap);
      }
      oshell *s = &b->sh[b->nshell++];
      memset(s, 0, sizeof *s);
      s->l = r->sh[k].l;
      s->pure = r->sh[k].pure;
      s->nprim = r->sh[k].nprim;
      memcpy(s->alpha, r->sh[k].alpha, sizeof s->alpha);
      memcpy(s->coef, r->sh[k].coef, sizeof s->coef);
      s->atom = a;
      for (int d = 0; d < 3; d++) s->O[d] = xyz[3 * a + d];
      normalize_shell(s);
    }
  }
  for (int i = 0; i < nrec; i++) free(recs[i].sh);
  int bf = 0;
  for (int i = 0; i < b->nshell; i++) {
    oshell *s = &b->sh[i];
    s->nfunc = s->pure ? 2 * s->l + 1 : NCART(s->l);
    s->bf = bf;
    bf += s->nfunc;
    if (s->nprim > b->maxprim) b->maxprim = s->nprim;
    if (s->l > b->maxl) b->maxl = s->l;
  }
  b->nbf = bf;
  compute_pairs(b);
  return b;
}

void oracle_basis_destroy(oracle_basis *b) {
  if (!b) return;
  free(b->sh);
  free(b->sig_off);
  free(b->sig_s2);
  free(b->schwarz);
  free(b);
}
int oracle_nbasis(const oracle_basis *b) { return b->nbf; }
int oracle_nshell(const oracle_basis *b) { return b->nshell; }
int oracle_maxprim(const oracle_basis *b) { return b->maxprim; }
int oracle_maxl(const oracle_basis *b) { return b->maxl; }
void oracle_shell_info(const oracle_basis *b, int s, int *l, int *pure,
                       int *nprim, int *bf, int *atom) {
  const oshell *x = &b->sh[s];
  *l = x->l; *pure = x->pure; *nprim = x->nprim; *bf = x->bf; *atom = x->atom;
}
void oracle_shell_prims(const oracle_basis *b, int s, double *alpha,
                        double *coef, double *origin) {
  const oshell *x = &b->sh[s];
  for (int p = 0; p < x->nprim; p++) { alpha[p] = x->alpha[p]; coef[p] = x->coef[p]; }
  for (int d = 0; d < 3; d++) origin[d] = x->O[d];
}

/* ------------------------------------------------------------------ */
/* Cartesian -> pure transforms on one index of a row-major tensor       */
/* ------------------------------------------------------------------ */
/* in: [pre][ncart][post] -> out: [pre][nsph][post] */
static void sph_index(int l, int pre, int post, const double *in, double *out) {
  int nc = NCART(l), ns = 2 * l + 1;
  for (int a = 0; a < pre; a++)
    for (int m = 0; m < ns; m++)
      for (int p = 0; p < post; p++) {
        double s = 0.0;
        for (int c = 0; c < nc; c++) s += SPH[l][m][c] * in[((size_t)a * nc + c) * post + p];
        out[((size_t)a * ns + m) * post + p] = s;
      }
}

/* ------------------------------------------------------------------ */
/* overlap (Obara-Saika)                                               */
/* ------------------------------------------------------------------ */
static void overlap_shell(const oshell *A, const oshell *B, double *out) {
  int la = A->l, lb = B->l;
  int nca = NCART(la), ncb = NCART(lb);
  double cart[NCART(MAXL) * NCART(MAXL)];
  memset(cart, 0, sizeof(double) * nca * ncb);
  double AB2 = 0;
  for (int d = 0; d < 3; d++) AB2 += (A->O[d] - B->O[d]) * (A->O[d] - B->O[d]);
  for (int p = 0; p < A->nprim; p++)
    for (int q = 0; q < B->nprim; q++) {
      double a = A->alpha[p], bb = B->alpha[q], g = a + bb;
      double K = A->coef[p] * B->coef[q] * exp(-a * bb / g * AB2) * pow(M_PI / g, 1.5);
      double S[3][MAXL + 2][MAXL + 2];
      for (int d = 0; d < 3; d++) {
        double P = (a * A->O[d] + bb * B->O[d]) / g;
        double PA = P - A->O[d], PB = P - B->O[d];
        S[d][0][0] = 1.0;
        for (int i = 0; i <= la; i++)
          for (int j = 0; j <= lb; j++) {
            if (i == 0 && j == 0) continue;
            double v;
            if (i > 0) {
              v = PA * S[d][i - 1][j];
              if (i > 1) v += (i - 1) / (2 * g) * S[d][i - 2][j];
              if (j > 0) v += j / (2 * g) * S[d][i - 1][j - 1];
            } else {
              v = PB * S[d][i][j - 1];
              if (j > 1) v += (j - 1) / (2 * g) * S[d][i][j - 2];
            }
            S[d][i][j] = v;
          }
      }
      for (int x = 0; x < nca; x++)
        for (int y = 0; y < ncb; y++) {
          ijk_t u = CART[NOFF(la) + x], w = CART[NOFF(lb) + y];
          cart[x * ncb + y] += K * S[0][u.x][w.x] * S[1][u.y][w.y] * S[2][u.z][w.z];
        }
    }
  double tmp[NCART(MAXL) * NCART(MAXL)];
  int na = A->pure ? 2 * la + 1 : nca;
  if (A->pure) sph_index(la, 1, ncb, cart, tmp); else memcpy(tmp, cart, sizeof(double) * nca * ncb);
  if (B->pure) sph_index(lb, na, 1, tmp, out); else memcpy(out, tmp, sizeof(double) * na * ncb);
}

void oracle_overlap(const oracle_basis *b, double *S) {
  int NB = b->nbf;
  double blk[NCART(MAXL) * NCART(MAXL)];
  for (int s1 = 0; s1 < b->nshell; s1++)
    for (int s2 = 0; s2 < b->nshell; s2++) {
      const oshell *A = &b->sh[s1], *B = &b->sh[s2];
      overlap_shell(A, B, blk);
      for (int i = 0; i < A->nfunc; i++)
        for (int j = 0; j < B->nfunc; j++)
          S[(A->bf + i) + (size_t)(B->bf + j) * NB] = blk[i * B->nfunc + j];
    }
}

/* significant pairs: ShellPairData::computeShellPairs
 * (src/basisset/basisset.cxx:652-705): s2 <= s1 kept if on the same centre
 * or if the 2-norm of the overlap block is >= 1e-12. */
static void compute_pairs(oracle_basis *b) {
  int ns = b->nshell;
  b->sig_off = malloc(sizeof(int) * (ns + 1));
  b->sig_s2 = malloc(sizeof(int) * (size_t)ns * (ns + 1) / 2 + 1);
  double blk[NCART(MAXL) * NCART(MAXL)];
  int n = 0;
  for (int s1 = 0; s1 < ns; s1++) {
    b->sig_off[s1] = n;
    for (int s2 = 0; s2 <= s1; s2++) {
      const oshell *A = &b->sh[s1], *B = &b->sh[s2];
      int same = fabs(A->O[0] - B->O[0]) < 1e-5 && fabs(A->O[1] - B->O[1]) < 1e-5 &&
                 fabs(A->O[2] - B->O[2]) < 1e-5;
      int sig = same;
      if (!same) {
        overlap_shell(A, B, blk);
        double s = 0;
        for (int k = 0; k < A->nfunc * B->nfunc; k++) s += blk[k] * blk[k];
        sig = sqrt(s) >= 1e-12;
      }
      if (sig) b->sig_s2[n++] = s2;
    }
  }
  b->sig_off[ns] = n;
  b->npairs = n;
}

int oracle_shell_pairs(const oracle_basis *b, int *s1, int *s2) {
  if (s1 && s2)
    for (int a = 0; a < b->nshell; a++)
      for (int k = b->sig_off[a]; k < b->sig_off[a + 1]; k++) {
        *s1++ = a;
        *s2++ = b->sig_s2[k];
      }
  return b->npairs;
}

/* ------------------------------------------------------------------ */
/* ERIs: Obara-Saika VRR + Head-Gordon-Pople HRR                        */
/* ------------------------------------------------------------------ */
typedef struct {
  double *V;    /* [nE][nF][Lt+1] */
  double *C;    /* contracted [e in la..Le][f in lc..Lf] */
  double *H1, *H2, *H3, *T1, *T2; /* HRR / transform scratch */
  double *O;    /* a quartet in the evaluation order (eri_shell) */
} eri_work;

static size_t work_elems(int maxl) {
  int Le = 2 * maxl;
  size_t nE = NOFF(Le + 1);
  return nE * nE * (2 * Le + 1);
}
static void work_alloc(eri_work *w, int maxl) {
  size_t nE = NOFF(2 * maxl + 1);
  w->V = malloc(sizeof(double) * work_elems(maxl));
  w->C = malloc(sizeof(double) * nE * nE);
  /* largest HRR level table: (#a) x (#b at that level) x (#spectators) */
  size_t nc = NCART(maxl), spec = nc * nc > nE ? nc * nc : nE;
  size_t big = nE * nc * spec + 16;
  if (big < nE * nE) big = nE * nE;
  w->H1 = malloc(sizeof(double) * big);
  w->H2 = malloc(sizeof(double) * big);
  w->H3 = malloc(sizeof(double) * big);
  w->T1 = malloc(sizeof(double) * big);
  w->T2 = malloc(sizeof(double) * big);
  w->O = malloc(sizeof(double) * nc * nc * nc * nc);
}
static void work_free(eri_work *w) {
  free(w->V); free(w->C); free(w->H1); free(w->H2); free(w->H3); free(w->T1); free(w->T2); free(w->O);
}

static inline int first_dir(ijk_t e) { return e.x ? 0 : (e.y ? 1 : 2); }
static inline int comp(ijk_t e, int d) { return d == 0 ? e.x : (d == 1 ? e.y : e.z); }
static inline int cidx(int x, int y, int z) { return NOFF(x + y + z) + cart_pos(x, y, z); }
static inline int cidx_dec(ijk_t e, int d, int k) {
  if (d == 0) return cidx(e.x - k, e.y, e.z);
  if (d == 1) return cidx(e.x, e.y - k, e.z);
  return cidx(e.x, e.y, e.z - k);
}
static inline int cidx_inc(ijk_t e, int d) {
  if (d == 0) return cidx(e.x + 1, e.y, e.z);
  if (d == 1) return cidx(e.x, e.y + 1, e.z);
  return cidx(e.x, e.y, e.z + 1);
}

/* HRR on the first index pair: in[e][spec], e over combined Cartesian
 * indices with l1 <= |e| <= l1+l2 (offset NOFF(l1)); out[a][b][spec]. */
static void hrr(int l1, int l2, const double AB[3], int nspec, const double *in,
                double *out, double *t1, double *t2) {
  int base = NOFF(l1);
  int nE = NOFF(l1 + l2 + 1) - base;
  /* level 0: cur[e][b0][spec], b0 = (0,0,0) */
  double *cur = t1, *nxt = t2;
  memcpy(cur, in, sizeof(double) * nE * nspec);
  for (int Lb = 1; Lb <= l2; Lb++) {
    int nbp = NCART(Lb - 1), nb = NCART(Lb);
    int amax = l1 + l2 - Lb;
    int nA = NOFF(amax + 1) - base;
    for (int ai = 0; ai < nA; ai++) {
      ijk_t a = CART[base + ai];
      for (int bi = 0; bi < nb; bi++) {
        ijk_t bb = CART[NOFF(Lb) + bi];
        int d = first_dir(bb);
        int bprev = cidx_dec(bb, d, 1) - NOFF(Lb - 1);
        int aup = cidx_inc(a, d) - base;
        double ab = AB[d];
        double *o = nxt + ((size_t)ai * nb + bi) * nspec;
        const double *p1 = cur + ((size_t)aup * nbp + bprev) * nspec;
        const double *p2 = cur + ((size_t)ai * nbp + bprev) * nspec;
        for (int s = 0; s < nspec; s++) o[s] = p1[s] + ab * p2[s];
      }
    }
    double *t = cur; cur = nxt; nxt = t;
  }
  /* cur: [a over levels l1..l1 (first NCART(l1))][b in l2][spec] */
  memcpy(out, cur, sizeof(double) * NCART(l1) * NCART(l2) * nspec);
}

static int eri_shell_p(const oshell *A, const oshell *B, const oshell *Cc, const oshell *D,
                       double *out, eri_work *w) {
  int la = A->l, lb = B->l, lc = Cc->l, ld = D->l;
  int Le = la + lb, Lf = lc + ld, Lt = Le + Lf;
  int nE = NOFF(Le + 1), nF = NOFF(Lf + 1), nM = Lt + 1;
  int e0 = NOFF(la), f0 = NOFF(lc);
  int nEc = nE - e0, nFc = nF - f0;
  double *V = w->V, *C = w->C;
  memset(C, 0, sizeof(double) * nEc * nFc);
  double AB[3], CD[3], AB2 = 0, CD2 = 0;
  for (int d = 0; d < 3; d++) {
    AB[d] = A->O[d] - B->O[d];
    CD[d] = Cc->O[d] - D->O[d];
    AB2 += AB[d] * AB[d];
    CD2 += CD[d] * CD[d];
  }
  const double pi52 = 2.0 * pow(M_PI, 2.5);
  double F[4 * MAXL + 2];
#define VI(e, f, m) V[((size_t)(e) * nF + (f)) * nM + (m)]
  for (int pa = 0; pa < A->nprim; pa++)
    for (int pb = 0; pb < B->nprim; pb++) {
      double a = A->alpha[pa], b = B->alpha[pb], p = a + b;
      double Kab = A->coef[pa] * B->coef[pb] * exp(-a * b / p * AB2);
      double P[3], PA[3];
      for (int d = 0; d < 3; d++) {
        P[d] = (a * A->O[d] + b * B->O[d]) / p;
        PA[d] = P[d] - A->O[d];
      }
      for (int pc = 0; pc < Cc->nprim; pc++)
        for (int pd = 0; pd < D->nprim; pd++) {
          double c = Cc->alpha[pc], dd = D->alpha[pd], q = c + dd;
          double Kcd = Cc->coef[pc] * D->coef[pd] * exp(-c * dd / q * CD2);
          double Q[3], QC[3], W[3], WP[3], WQ[3], PQ2 = 0;
          for (int d = 0; d < 3; d++) {
            Q[d] = (c * Cc->O[d] + dd * D->O[d]) / q;
            QC[d] = Q[d] - Cc->O[d];
            W[d] = (p * P[d] + q * Q[d]) / (p + q);
            WP[d] = W[d] - P[d];
            WQ[d] = W[d] - Q[d];
            PQ2 += (P[d] - Q[d]) * (P[d] - Q[d]);
          }
          double rho = p * q / (p + q);
          double pref = Kab * Kcd * pi52 / (p * q * sqrt(p + q));
          oracle_boys(Lt, rho * PQ2, F);
          for (int m = 0; m <= Lt; m++) VI(0, 0, m) = pref * F[m];
          /* bra VRR */
          for (int E = 1; E <= Le; E++)
            for (int ei = NOFF(E); ei < NOFF(E + 1); ei++) {
              ijk_t e = CART[ei];
              int d = first_dir(e);
              int e1 = cidx_dec(e, d, 1);
              int n1 = comp(e, d) - 1;
              int e2 = n1 > 0 ? cidx_dec(e, d, 2) : -1;
              for (int m = 0; m <= Lt - E; m++) {
                double v = PA[d] * VI(e1, 0, m) + WP[d] * VI(e1, 0, m + 1);
                if (n1 > 0) v += n1 / (2 * p) * (VI(e2, 0, m) - rho / p * VI(e2, 0, m + 1));
                VI(ei, 0, m) = v;
              }
            }
          /* ket VRR */
          for (int Fl = 1; Fl <= Lf; Fl++)
            for (int fi = NOFF(Fl); fi < NOFF(Fl + 1); fi++) {
              ijk_t f = CART[fi];
              int d = first_dir(f);
              int f1 = cidx_dec(f, d, 1);
              int n1 = comp(f, d) - 1;
              int f2 = n1 > 0 ? cidx_dec(f, d, 2) : -1;
              for (int ei = 0; ei < nE; ei++) {
                ijk_t e = CART[ei];
                int E = e.x + e.y + e.z;
                int ne = comp(e, d);
                int em = ne > 0 ? cidx_dec(e, d, 1) : -1;
                for (int m = 0; m <= Lt - E - Fl; m++) {
                  double v = QC[d] * VI(ei, f1, m) + WQ[d] * VI(ei, f1, m + 1);
                  if (n1 > 0) v += n1 / (2 * q) * (VI(ei, f2, m) - rho / q * VI(ei, f2, m + 1));
                  if (ne > 0) v += ne / (2 * (p + q)) * VI(em, f1, m + 1);
                  VI(ei, fi, m) = v;
                }
              }
            }
          for (int ei = e0; ei < nE; ei++)
            for (int fi = f0; fi < nF; fi++) C[(ei - e0) * nFc + (fi - f0)] += VI(ei, fi, 0);
        }
    }
#undef VI
  /* bra HRR: C[e][f] -> H1[a][b][f] */
  int nca = NCART(la), ncb = NCART(lb), ncc = NCART(lc), ncd = NCART(ld);
  hrr(la, lb, AB, nFc, C, w->H1, w->T1, w->T2);
  /* ket HRR per (a,b): need [f][spec] layout; transpose into H2[f][ab] */
  int nab = nca * ncb;
  for (int ab = 0; ab < nab; ab++)
    for (int f = 0; f < nFc; f++) w->H2[(size_t)f * nab + ab] = w->H1[(size_t)ab * nFc + f];
  hrr(lc, ld, CD, nab, w->H2, w->H3, w->T1, w->T2);
  /* H3[c][d][ab] -> cart[ab][c][d] */
  int ncd_ = ncc * ncd;
  double *cart = w->H1;
  for (int ab = 0; ab < nab; ab++)
    for (int cd = 0; cd < ncd_; cd++) cart[(size_t)ab * ncd_ + cd] = w->H3[(size_t)cd * nab + ab];
  /* pure transforms, index by index */
  int n1 = nca, n2 = ncb, n3 = ncc, n4 = ncd;
  double *src = cart, *dst = w->H2;
  if (A->pure) { sph_index(la, 1, n2 * n3 * n4, src, dst); n1 = 2 * la + 1; double *t = src; src = dst; dst = t; }
  if (B->pure) { sph_index(lb, n1, n3 * n4, src, dst); n2 = 2 * lb + 1; double *t = src; src = dst; dst = t; }
  if (Cc->pure) { sph_index(lc, n1 * n2, n4, src, dst); n3 = 2 * lc + 1; double *t = src; src = dst; dst = t; }
  if (D->pure) { sph_index(ld, n1 * n2 * n3, 1, src, dst); n4 = 2 * ld + 1; double *t = src; src = dst; dst = t; }
  int n = n1 * n2 * n3 * n4;
  memcpy(out, src, sizeof(double) * n);
  return n;
}

/* The HRR moves angular momentum from the first shell of a pair onto the
 * second over the distance between them; moving the smaller one loses less
 * to cancellation (a g shell second in a pair costs 1e-10 relative at O-H
 * distances), so each pair is evaluated with its higher-l shell first and
 * the quartet is transposed back to the order asked for. */
static int eri_oriented(const oshell *S1, const oshell *S2, const oshell *S3, const oshell *S4,
                        double *out, eri_work *w) {
  const int swb = S1->l < S2->l, swk = S3->l < S4->l;
  if (!swb && !swk) return eri_shell_p(S1, S2, S3, S4, out, w);
  const oshell *A = swb ? S2 : S1, *B = swb ? S1 : S2, *C = swk ? S4 : S3, *D = swk ? S3 : S4;
  const int n = eri_shell_p(A, B, C, D, w->O, w);
  const int n1 = S1->nfunc, n2 = S2->nfunc, n3 = S3->nfunc, n4 = S4->nfunc;
  const int nb = B->nfunc, nc = C->nfunc, nd = D->nfunc;
  for (int i1 = 0; i1 < n1; i1++)
    for (int i2 = 0; i2 < n2; i2++)
      for (int i3 = 0; i3 < n3; i3++)
        for (int i4 = 0; i4 < n4; i4++) {
          const int ia = swb ? i2 : i1, ib = swb ? i1 : i2, ic = swk ? i4 : i3, id = swk ? i3 : i4;
          out[(((size_t)i1 * n2 + i2) * n3 + i3) * n4 + i4] = w->O[(((size_t)ia * nb + ib) * nc + ic) * nd + id];
        }
  return n;
}

static int eri_shell(const oracle_basis *bs, int s1, int s2, int s3, int s4,
                     double *out, eri_work *w) {
  return eri_oriented(&bs->sh[s1], &bs->sh[s2], &bs->sh[s3], &bs->sh[s4], out, w);
}

int oracle_eri_quartet(const oracle_basis *b, int s1, int s2, int s3, int s4,
                       double *out) {
  build_tables();
  eri_work w;
  work_alloc(&w, b->maxl);
  int n = eri_shell(b, s1, s2, s3, s4, out, &w);
  work_free(&w);
  return n;
}

void oracle_eri_full(const oracle_basis *b, double *out) {
  int NB = b->nbf, ns = b->nshell;
  size_t NB2 = (size_t)NB * NB;
#pragma omp parallel
  {
    eri_work w;
    work_alloc(&w, b->maxl);
    size_t mx = 1;
    for (int i = 0; i < 4; i++) mx *= NCART(b->maxl);
    double *buf = malloc(sizeof(double) * mx);
#pragma omp for schedule(dynamic) collapse(2)
    for (int s1 = 0; s1 < ns; s1++)
      for (int s2 = 0; s2 < ns; s2++)
        for (int s3 = 0; s3 < ns; s3++)
          for (int s4 = 0; s4 < ns; s4++) {
            const oshell *A = &b->sh[s1], *B = &b->sh[s2], *C = &b->sh[s3], *D = &b->sh[s4];
            eri_shell(b, s1, s2, s3, s4, buf, &w);
            int ijkl = 0;
            for (int i = 0; i < A->nfunc; i++)
              for (int j = 0; j < B->nfunc; j++)
                for (int k = 0; k < C->nfunc; k++)
                  for (int l = 0; l < D->nfunc; l++, ijkl++)
                    out[((size_t)(A->bf + i) * NB + (B->bf + j)) * NB2 + (size_t)(C->bf + k) * NB +
                        (D->bf + l)] = buf[ijkl];
          }
    free(buf);
    work_free(&w);
  }
}

/* ------------------------------------------------------------------ */
/* Schwarz: DirectTPI::computeSchwarz                                   */
/* (src/particleintegrals/aointegrals/aointegrals_utilities.cxx:53-162) */
/* Q(s1,s2) = sqrt(|| [(ij|ij)]_{n1 x n2} ||_inf), mirrored.            */
/* ------------------------------------------------------------------ */
void oracle_schwarz(const oracle_basis *b, double *Q) {
  int ns = b->nshell;
#pragma omp parallel
  {
    eri_work w;
    work_alloc(&w, b->maxl);
    size_t mx = 1;
    for (int i = 0; i < 4; i++) mx *= NCART(b->maxl);
    double *buf = malloc(sizeof(double) * mx);
#pragma omp for schedule(dynamic)
    for (int s1 = 0; s1 < ns; s1++)
      for (int s2 = 0; s2 <= s1; s2++) {
        int n1 = b->sh[s1].nfunc, n2 = b->sh[s2].nfunc;
        eri_shell(b, s1, s2, s1, s2, buf, &w);
        double nrm = 0.0;
        for (int i = 0; i < n1; i++) {
          double row = 0.0;
          for (int j = 0; j < n2; j++) {
            size_t ij = (size_t)i * n2 + j;
            row += fabs(buf[ij * n1 * n2 + ij]);
          }
          if (row > nrm) nrm = row;
        }
        Q[s1 + (size_t)s2 * ns] = sqrt(nrm);
        Q[s2 + (size_t)s1 * ns] = sqrt(nrm);
      }
    free(buf);
    work_free(&w);
  }
}

/* ------------------------------------------------------------------ */
/* ShellBlockNorm (direct.hpp:73-105): max absolute row sum per block   */
/* ------------------------------------------------------------------ */
static void block_norms(const oracle_basis *b, const void *X, int is_complex, double *B) {
  int ns = b->nshell, NB = b->nbf;
  for (int s1 = 0; s1 < ns; s1++)
    for (int s2 = 0; s2 < ns; s2++) {
      const oshell *A = &b->sh[s1], *S = &b->sh[s2];
      double nrm = 0.0;
      for (int i = 0; i < A->nfunc; i++) {
        double row = 0.0;
        for (int j = 0; j < S->nfunc; j++) {
          size_t idx = (A->bf + i) + (size_t)(S->bf + j) * NB;
          row += is_complex ? cabs(((const double complex *)X)[idx]) : fabs(((const double *)X)[idx]);
        }
        if (row > nrm) nrm = row;
      }
      B[s1 + (size_t)s2 * ns] = nrm;
    }
}

/* ------------------------------------------------------------------ */
/* digestion, restating direct.hpp:1858-2002 for real and complex X     */
/* ------------------------------------------------------------------ */
#define DIGEST_BODY(T, CONJ, REALPART)                                                      \
  {                                                                                         \
    const T *X = (const T *)Xv;                                                             \
    T *AX = (T *)AXv;                                                                       \
    size_t ijkl = 0;                                                                        \
    if (her) {                                                                              \
      if (type == 0) {                                                                      \
        for (int i = 0, bf1 = b1s; i < n1; i++, bf1++)                                      \
          for (int j = 0, bf2 = b2s; j < n2; j++, bf2++) {                                  \
            size_t p1 = bf1 + (size_t)NB * bf2;                                             \
            double X1 = REALPART(X[p1]);                                                    \
            double *A1 = (double *)&AX[p1];                                                 \
            for (int k = 0, bf3 = b3s; k < n3; k++, bf3++)                                  \
              for (int l = 0, bf4 = b4s; l < n4; l++, bf4++, ijkl++) {                      \
                *A1 += REALPART(X[bf4 + (size_t)bf3 * NB]) * ERI[ijkl];                       \
                *(double *)&AX[bf4 + (size_t)bf3 * NB] += X1 * ERI[ijkl];                     \
              }                                                                             \
          }                                                                                 \
      } else {                                                                              \
        for (int i = 0, bf1 = b1s; i < n1; i++, bf1++)                                      \
          for (int j = 0, bf2 = b2s; j < n2; j++, bf2++)                                    \
            for (int k = 0, bf3 = b3s; k < n3; k++, bf3++) {                                \
              size_t q1 = bf1 + (size_t)bf3 * NB, q2 = bf2 + (size_t)bf3 * NB;              \
              T T1 = 0.5 * CONJ(X[q1]), T2 = 0.5 * CONJ(X[q2]);                             \
              for (int l = 0, bf4 = b4s; l < n4; l++, bf4++, ijkl++) {                      \
                double v = ERI[ijkl];                                                         \
                AX[q1] += 0.5 * CONJ(X[bf4 + (size_t)NB * bf2]) * v;                        \
                AX[bf4 + (size_t)bf2 * NB] += T1 * v;                                       \
                AX[bf4 + (size_t)bf1 * NB] += T2 * v;                                       \
                AX[q2] += 0.5 * CONJ(X[bf4 + (size_t)NB * bf1]) * v;                        \
              }                                                                             \
            }                                                                               \
      }                                                                                     \
    } else {                                                                                \
      if (type == 0) {                                                                      \
        for (int i = 0, bf1 = b1s; i < n1; i++, bf1++)                                      \
          for (int j = 0, bf2 = b2s; j < n2; j++, bf2++) {                                  \
            size_t p1 = bf1 + (size_t)NB * bf2, p2 = bf2 + (size_t)NB * bf1;                \
            T T1 = X[p1], T2 = X[p2];                                                       \
            for (int k = 0, bf3 = b3s; k < n3; k++, bf3++)                                  \
              for (int l = 0, bf4 = b4s; l < n4; l++, bf4++, ijkl++) {                      \
                double v = ERI[ijkl];                                                         \
                T x34 = 0.5 * (X[bf4 + (size_t)bf3 * NB] + X[bf3 + (size_t)bf4 * NB]);      \
                AX[p1] += x34 * v;                                                          \
                AX[bf3 + (size_t)bf4 * NB] += 0.5 * (T2 + T1) * v;                          \
                AX[p2] += x34 * v;                                                          \
                AX[bf4 + (size_t)bf3 * NB] += 0.5 * (T2 + T1) * v;                          \
              }                                                                             \
          }                                                                                 \
      } else {                                                                              \
        for (int i = 0, bf1 = b1s; i < n1; i++, bf1++)                                      \
          for (int j = 0, bf2 = b2s; j < n2; j++, bf2++)                                    \
            for (int k = 0, bf3 = b3s; k < n3; k++, bf3++) {                                \
              size_t q1 = bf1 + (size_t)bf3 * NB, q2 = bf2 + (size_t)bf3 * NB;              \
              size_t q3 = bf3 + (size_t)bf1 * NB, q4 = bf3 + (size_t)bf2 * NB;              \
              T T1 = 0.5 * X[q1], T2 = 0.5 * X[q2], T3 = 0.5 * X[q3], T4 = 0.5 * X[q4];     \
              for (int l = 0, bf4 = b4s; l < n4; l++, bf4++, ijkl++) {                      \
                double v = ERI[ijkl];                                                         \
                AX[q3] += 0.5 * X[bf4 + (size_t)NB * bf2] * v;                              \
                AX[bf4 + (size_t)bf2 * NB] += T3 * v;                                       \
                AX[bf4 + (size_t)bf1 * NB] += T4 * v;                                       \
                AX[q4] += 0.5 * X[bf4 + (size_t)NB * bf1] * v;                              \
                AX[q1] += 0.5 * X[bf2 + (size_t)NB * bf4] * v;                              \
                AX[bf2 + (size_t)bf4 * NB] += T1 * v;                                       \
                AX[bf1 + (size_t)bf4 * NB] += T2 * v;                                       \
                AX[q2] += 0.5 * X[bf1 + (size_t)NB * bf4] * v;                              \
              }                                                                             \
            }                                                                               \
      }                                                                                     \
    }                                                                                       \
  }

#define RID(x) (x)
#define RPART_R(x) (x)
static void digest_real(const void *Xv, void *AXv, int her, int type, int NB,
                        int b1s, int b2s, int b3s, int b4s, int n1, int n2, int n3,
                        int n4, const double *ERI) DIGEST_BODY(double, RID, RPART_R)

#define CCONJ(x) conj(x)
#define RPART_C(x) creal(x)
static void digest_complex(const void *Xv, void *AXv, int her, int type, int NB,
                           int b1s, int b2s, int b3s, int b4s, int n1, int n2,
                           int n3, int n4, const double *ERI)
    DIGEST_BODY(double complex, CCONJ, RPART_C)

/* ------------------------------------------------------------------ */
/* GTODirectTPIContraction::directScaffoldNew (direct.hpp:1412-2123)    */
/* ------------------------------------------------------------------ */
int oracle_two_body_contract(oracle_basis *b, double thresh, int screen,
                             int nmat, const oracle_contraction *cl,
                             int is_complex, int nthreads) {
  build_tables();
  if (nmat <= 0) { set_err("empty contraction list"); return 1; }
  int ns = b->nshell, NB = b->nbf;
  size_t NB2 = (size_t)NB * NB, ns2 = (size_t)ns * ns;
  size_t esz = is_complex ? sizeof(double complex) : sizeof(double);
  if (nthreads <= 0) nthreads = omp_get_max_threads();

  if (!b->schwarz) {
    b->schwarz = malloc(sizeof(double) * ns2);
    oracle_schwarz(b, b->schwarz);
  }
  const double *Q = b->schwarz;

  /* per-matrix shell-block norms, symmetrised for non-Hermitian X, and
   * their elementwise max over the matrices (direct.hpp:1515-1561) */
  int off = nmat == 1 ? 0 : 1;
  double *Braw = calloc((size_t)(nmat + off) * ns2, sizeof(double));
  double *Bg = Braw;
  for (int m = 0; m < nmat; m++) {
    double *Bm = Braw + (size_t)(m + off) * ns2;
    block_norms(b, cl[m].X, is_complex, Bm);
    if (!cl[m].her)
      for (int k = 0; k < ns; k++)
        for (int l = 0; l < k; l++) {
          double mx = fmax(Bm[k + (size_t)l * ns], Bm[l + (size_t)k * ns]);
          Bm[k + (size_t)l * ns] = mx;
          Bm[l + (size_t)k * ns] = mx;
        }
  }
  if (nmat != 1)
    for (size_t i = 0; i < ns2; i++)
      for (int m = 0; m < nmat; m++) Bg[i] = fmax(Bg[i], Braw[(size_t)(m + 1) * ns2 + i]);
#define BM(m, x, y) Braw[(size_t)((m) + off) * ns2 + (x) + (size_t)(y) * ns]

  /* per-thread accumulators (direct.hpp:1495-1513) */
  char *AXraw = NULL;
  if (nthreads != 1) AXraw = calloc((size_t)nthreads * nmat * NB2, esz);
  long long n_comp = 0, n_skip = 0;

#pragma omp parallel num_threads(nthreads) reduction(+ : n_comp, n_skip)
  {
    int tid = omp_get_thread_num();
    eri_work w;
    work_alloc(&w, b->maxl);
    size_t mx = 1;
    for (int i = 0; i < 4; i++) mx *= NCART(b->maxl);
    double *buf = malloc(sizeof(double) * mx);
    int *cmat = malloc(sizeof(int) * nmat);
    double *sh123m = malloc(sizeof(double) * nmat);
    void **AXl = malloc(sizeof(void *) * nmat);
    for (int m = 0; m < nmat; m++)
      AXl[m] = nthreads == 1 ? cl[m].AX : (void *)(AXraw + ((size_t)tid * nmat + m) * NB2 * esz);

    size_t s12 = 0;
    for (int s1 = 0; s1 < ns; s1++) {
      const oshell *S1 = &b->sh[s1];
      for (int k2 = b->sig_off[s1]; k2 < b->sig_off[s1 + 1]; k2++) {
        int s2 = b->sig_s2[k2];
        const oshell *S2 = &b->sh[s2];
        if ((s12++) % (size_t)nthreads != (size_t)tid) continue;
        double s12_deg = (s1 == s2) ? 1.0 : 2.0;
        double shz12 = 0, shMax12 = 0;
        if (screen) {
          shz12 = Q[s1 + (size_t)s2 * ns];
          shMax12 = Bg[s1 + (size_t)s2 * ns];
        }
        for (int s3 = 0; s3 <= s1; s3++) {
          const oshell *S3 = &b->sh[s3];
          double shMax123 = 0;
          if (screen) {
            shMax123 = fmax(Bg[s1 + (size_t)s3 * ns], Bg[s2 + (size_t)s3 * ns]);
            shMax123 = fmax(shMax123, shMax12);
            if (nmat != 1)
              for (int m = 0; m < nmat; m++)
                sh123m[m] = fmax(BM(m, s1, s2), fmax(BM(m, s1, s3), BM(m, s2, s3)));
          }
          int s4max = (s1 == s3) ? s2 : s3;
          for (int k4 = b->sig_off[s3]; k4 < b->sig_off[s3 + 1]; k4++) {
            int s4 = b->sig_s2[k4];
            if (s4 > s4max) break;
            const oshell *S4 = &b->sh[s4];
            int ncon = nmat;
            for (int m = 0; m < nmat; m++) cmat[m] = m;
            if (screen) {
              double shMax = Bg[s3 + (size_t)s4 * ns];
              shMax = fmax(shMax, fmax(Bg[s1 + (size_t)s4 * ns], Bg[s2 + (size_t)s4 * ns]));
              shMax = fmax(shMax, shMax123);
              if (shMax * shz12 * Q[s3 + (size_t)s4 * ns] < thresh) { n_skip++; continue; }
              if (nmat != 1) {
                ncon = 0;
                for (int m = 0; m < nmat; m++) {
                  double sm = BM(m, s3, s4);
                  sm = fmax(sm, fmax(BM(m, s1, s4), BM(m, s2, s4)));
                  sm = fmax(sm, sh123m[m]);
                  if (!(sm * shz12 * Q[s3 + (size_t)s4 * ns] < thresh)) cmat[ncon++] = m;
                }
              }
            }
            double s34_deg = (s3 == s4) ? 1.0 : 2.0;
            double s1234_deg = (s1 == s3) ? (s2 == s4 ? 1.0 : 2.0) : 2.0;
            double scale = 0.5 * s12_deg * s34_deg * s1234_deg;
            int n = eri_shell(b, s1, s2, s3, s4, buf, &w);
            for (int i = 0; i < n; i++) buf[i] *= scale;
            n_comp++;
            for (int ic = 0; ic < ncon; ic++) {
              int m = cmat[ic];
              if (is_complex)
                digest_complex(cl[m].X, AXl[m], cl[m].her, cl[m].type, NB, S1->bf, S2->bf,
                               S3->bf, S4->bf, S1->nfunc, S2->nfunc, S3->nfunc, S4->nfunc, buf);
              else
                digest_real(cl[m].X, AXl[m], cl[m].her, cl[m].type, NB, S1->bf, S2->bf,
                            S3->bf, S4->bf, S1->nfunc, S2->nfunc, S3->nfunc, S4->nfunc, buf);
            }
          }
        }
      }
    }
    free(AXl);
    free(sh123m);
    free(cmat);
    free(buf);
    work_free(&w);
  }
#undef BM

  /* thread reduction and Hermitian symmetrisation (direct.hpp:2036-2066) */
  for (int m = 0; m < nmat; m++)
    for (int t = 0; t < nthreads; t++) {
      void *src = nthreads == 1 ? cl[m].AX : (void *)(AXraw + ((size_t)t * nmat + m) * NB2 * esz);
      if (is_complex) {
        double complex *A = src, *O = cl[m].AX;
        if (cl[m].her) {
          double complex *scr = malloc(sizeof(double complex) * NB2);
          for (int j = 0; j < NB; j++)
            for (int i = 0; i < NB; i++)
              scr[i + (size_t)j * NB] = 0.5 * A[i + (size_t)j * NB] + 0.5 * conj(A[j + (size_t)i * NB]);
          for (size_t k = 0; k < NB2; k++) O[k] = nthreads != 1 ? O[k] + scr[k] : scr[k];
          free(scr);
        } else {
          for (size_t k = 0; k < NB2; k++) O[k] = nthreads != 1 ? O[k] + 0.5 * A[k] : 0.5 * O[k];
        }
      } else {
        double *A = src, *O = cl[m].AX;
        if (cl[m].her) {
          double *scr = malloc(sizeof(double) * NB2);
          for (int j = 0; j < NB; j++)
            for (int i = 0; i < NB; i++)
              scr[i + (size_t)j * NB] = 0.5 * A[i + (size_t)j * NB] + 0.5 * A[j + (size_t)i * NB];
          for (size_t k = 0; k < NB2; k++) O[k] = nthreads != 1 ? O[k] + scr[k] : scr[k];
          free(scr);
        } else {
          for (size_t k = 0; k < NB2; k++) O[k] = nthreads != 1 ? O[k] + 0.5 * A[k] : 0.5 * O[k];
        }
      }
    }
  free(AXraw);
  free(Braw);
  b->st_computed = n_comp;
  b->st_skipped = n_skip;
  return 0;
}

/* ------------------------------------------------------------------ */
/* DirectGradContraction::directScaffoldGrad (direct.hpp:2198-2580)     */
/* ------------------------------------------------------------------ */
/* A shell moved up or down one unit of l on the same primitives, Cartesian:
 * the derivative of a primitive x_A^a e^{-alpha r_A^2} along x is
 * 2 alpha x_A^{a+1} (...) - a x_A^{a-1} (...), so the up shell carries the
 * coefficients 2 alpha c and the down shell c. */
static void shifted_shell(const oshell *S, int up, oshell *V) {
  *V = *S;
  V->pure = 0;
  V->l = S->l + (up ? 1 : -1);
  V->nfunc = NCART(V->l);
  if (up)
    for (int i = 0; i < S->nprim; i++) V->coef[i] = 2.0 * S->alpha[i] * S->coef[i];
}

/* Derivatives of (s1 s2|s3 s4) with respect to centre k (0..3) along x, y
 * and z: out[d] in eri_shell's layout (pure indices where the shell is
 * pure).  scr holds three buffers of the largest quartet with one index
 * raised by one unit of l. */
static void eri_centre_deriv(const oracle_basis *bs, const int s[4], int k, double *out[3],
                             eri_work *w, double *scr, size_t scr_n) {
  const oshell *S[4] = {&bs->sh[s[0]], &bs->sh[s[1]], &bs->sh[s[2]], &bs->sh[s[3]]};
  const oshell *K = S[k];
  oshell up, dn;
  shifted_shell(K, 1, &up);
  const oshell *Q[4] = {S[0], S[1], S[2], S[3]};
  Q[k] = &up;
  double *P = scr, *M = scr + scr_n, *T = scr + 2 * scr_n;
  eri_oriented(Q[0], Q[1], Q[2], Q[3], P, w);
  if (K->l > 0) {
    shifted_shell(K, 0, &dn);
    Q[k] = &dn;
    eri_oriented(Q[0], Q[1], Q[2], Q[3], M, w);
  }
  int pre = 1, post = 1;
  for (int j = 0; j < k; j++) pre *= S[j]->nfunc;
  for (int j = k + 1; j < 4; j++) post *= S[j]->nfunc;
  const int l = K->l, nc = NCART(l), nu = NCART(l + 1), nd = l > 0 ? NCART(l - 1) : 0;
  for (int d = 0; d < 3; d++) {
    /* Cartesian derivative along d in T: [pre][nc][post] */
    for (int c = 0; c < nc; c++) {
      ijk_t e = CART[NOFF(l) + c];
      int iu = cidx_inc(e, d) - NOFF(l + 1);
      int n = comp(e, d);
      int id = n > 0 ? cidx_dec(e, d, 1) - NOFF(l - 1) : -1;
      for (int a = 0; a < pre; a++)
        for (int z = 0; z < post; z++) {
          double v = P[((size_t)a * nu + iu) * post + z];
          if (n > 0) v -= n * M[((size_t)a * nd + id) * post + z];
          T[((size_t)a * nc + c) * post + z] = v;
        }
    }
    if (K->pure)
      sph_index(l, pre, post, T, out[d]);
    else
      memcpy(out[d], T, sizeof(double) * (size_t)pre * nc * post);
  }
}

int oracle_grad_contract(oracle_basis *b, int nmat, const oracle_contraction *cl, int nthreads) {
  build_tables();
  if (nmat <= 0) { set_err("empty contraction list"); return 1; }
  const int ngrad = 3 * b->natom, NB = b->nbf, ns = b->nshell;
  for (int i = 0; i < ngrad * nmat; i++)
    if (!cl[i].her) { set_err("Nonhermetian NYI (directScaffoldGrad)"); return 1; }
  const size_t NB2 = (size_t)NB * NB;
  if (nthreads <= 0) nthreads = omp_get_max_threads();
  double *acc = calloc((size_t)nthreads * ngrad * nmat * NB2, sizeof(double));
  if (!acc) { set_err("out of memory"); return 1; }
  const int ml = b->maxl + 1;
  size_t scr_n = 1;
  for (int i = 0; i < 4; i++) scr_n *= NCART(ml);

#pragma omp parallel num_threads(nthreads)
  {
    const int tid = omp_get_thread_num();
    eri_work w;
    work_alloc(&w, ml);
    double *scr = malloc(sizeof(double) * 3 * scr_n);
    double *dbuf = malloc(sizeof(double) * 3 * scr_n);
    double *dI[3] = {dbuf, dbuf + scr_n, dbuf + 2 * scr_n};
    double *A = acc + (size_t)tid * ngrad * nmat * NB2;
    size_t s12 = 0;
    for (int s1 = 0; s1 < ns; s1++)
      for (int k2 = b->sig_off[s1]; k2 < b->sig_off[s1 + 1]; k2++) {
        const int s2 = b->sig_s2[k2];
        if ((s12++) % (size_t)nthreads != (size_t)tid) continue;
        for (int s3 = 0; s3 <= s1; s3++) {
          const int s4max = (s1 == s3) ? s2 : s3;
          for (int k4 = b->sig_off[s3]; k4 < b->sig_off[s3 + 1]; k4++) {
            const int s4 = b->sig_s2[k4];
            if (s4 > s4max) break;
            const int s[4] = {s1, s2, s3, s4};
            const oshell *S[4] = {&b->sh[s1], &b->sh[s2], &b->sh[s3], &b->sh[s4]};
            const double deg = (s1 == s2 ? 1.0 : 2.0) * (s3 == s4 ? 1.0 : 2.0) *
                               ((s1 == s3) ? (s2 == s4 ? 1.0 : 2.0) : 2.0);
            const size_t n = (size_t)S[0]->nfunc * S[1]->nfunc * S[2]->nfunc * S[3]->nfunc;
            /* the twelve derivative buffers, centre by centre (:2390-2405) */
            for (int k = 0; k < 4; k++) {
              eri_centre_deriv(b, s, k, dI, &w, scr, scr_n);
              for (int d = 0; d < 3; d++) {
                for (size_t i = 0; i < n; i++) dI[d][i] *= 0.5 * deg;
                const int g = S[k]->atom * 3 + d;
                for (int m = 0; m < nmat; m++)
                  digest_real(cl[(size_t)g * nmat + m].X, A + ((size_t)g * nmat + m) * NB2, 1,
                              cl[(size_t)g * nmat + m].type, NB, S[0]->bf, S[1]->bf, S[2]->bf, S[3]->bf,
                              S[0]->nfunc, S[1]->nfunc, S[2]->nfunc, S[3]->nfunc, dI[d]);
              }
            }
          }
        }
      }
    free(dbuf);
    free(scr);
    work_free(&w);
  }
  /* Hermitian symmetrisation per thread, added to AX (:2548-2570) */
  for (int gm = 0; gm < ngrad * nmat; gm++) {
    double *O = cl[gm].AX;
    for (int t = 0; t < nthreads; t++) {
      const double *At = acc + ((size_t)t * ngrad * nmat + gm) * NB2;
      for (int j = 0; j < NB; j++)
        for (int i = 0; i < NB; i++)
          O[i + (size_t)j * NB] += 0.5 * At[i + (size_t)j * NB] + 0.5 * At[j + (size_t)i * NB];
    }
  }
  free(acc);
  return 0;
}

void oracle_prepare(oracle_basis *b) {
  build_tables();
  if (!b->schwarz) {
    b->schwarz = malloc(sizeof(double) * (size_t)b->nshell * b->nshell);
    oracle_schwarz(b, b->schwarz);
  }
}

void oracle_last_stats(const oracle_basis *b, long long *nc, long long *nsk) {
  *nc = b->st_computed;
  *nsk = b->st_skipped;
}
