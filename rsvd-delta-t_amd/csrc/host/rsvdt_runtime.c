/*
 * Runtime of the C host layer: initialization, errors, the options table
 * that plays the role of PETSc's options database (filled from the command
 * line and from variables.yaml), Mat handles and the PETSc-binary reader and
 * writer (the MatLoad/MatView of SourceCode/LNSStructure.c:23-26,
 * ReadWeightMats.c:27-87, CreateRandomMat.c:24-27 and SaveResults.c:32-85).
 */
#define _POSIX_C_SOURCE 200809L
#include "rsvdt_internal.h"

#include <ctype.h>
#include <math.h>
#include <stdarg.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>
#include <sys/time.h>

rsvdt_runtime g_rsvdt;

/* ------------------------------------------------------------------ errors */
PetscErrorCode rsvdt_err(PetscErrorCode code, const char *fn, const char *fmt, ...)
{
    va_list ap;
    fprintf(stderr, "[rsvdt] error %d in %s: ", code, fn);
    va_start(ap, fmt);
    vfprintf(stderr, fmt, ap);
    va_end(ap);
    fprintf(stderr, "\n");
    fflush(stderr);
    return code;
}

int rsvdt_printf(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    int r = vprintf(fmt, ap);
    va_end(ap);
    fflush(stdout);
    return r;
}

PetscLogDouble RSVDT_Time(void)
{
    struct timeval tv;
    gettimeofday(&tv, NULL);
    return (double)tv.tv_sec + 1e-6 * (double)tv.tv_usec;
}

/* ------------------------------------------------------------------ options */
typedef struct { char *key; char *val; } opt_entry;
static opt_entry *g_opts = NULL;
static size_t g_nopts = 0, g_capopts = 0;

PetscErrorCode RSVDT_OptionsSetValue(const char *key, const char *value)
{
    if (!key) return 0;
    if (key[0] == '-') key++;
    for (size_t i = 0; i < g_nopts; i++)
        if (strcmp(g_opts[i].key, key) == 0) {
            free(g_opts[i].val);
            g_opts[i].val = strdup(value ? value : "");
            return 0;
        }
    if (g_nopts == g_capopts) {
        g_capopts = g_capopts ? 2 * g_capopts : 32;
        g_opts = realloc(g_opts, g_capopts * sizeof(opt_entry));
        if (!g_opts) return rsvdt_err(RSVDT_ERR_MEM, "RSVDT_OptionsSetValue", "out of memory");
    }
    g_opts[g_nopts].key = strdup(key);
    g_opts[g_nopts].val = strdup(value ? value : "");
    g_nopts++;
    return 0;
}

PetscErrorCode RSVDT_OptionsClear(void)
{
    for (size_t i = 0; i < g_nopts; i++) { free(g_opts[i].key); free(g_opts[i].val); }
    free(g_opts);
    g_opts = NULL; g_nopts = g_capopts = 0;
    return 0;
}

static const char *opt_lookup(const char *key)
{
    if (key[0] == '-') key++;
    for (size_t i = 0; i < g_nopts; i++)
        if (strcmp(g_opts[i].key, key) == 0) return g_opts[i].val;
    return NULL;
}

int rsvdt_opt_get_string(const char *key, char *out, size_t n)
{
    const char *v = opt_lookup(key);
    if (!v) return 0;
    snprintf(out, n, "%s", v);
    return 1;
}

int rsvdt_opt_get_real(const char *key, double *out, int *ok)
{
    const char *v = opt_lookup(key);
    *ok = 1;
    if (!v) return 0;
    char *end;
    double d = strtod(v, &end);
    if (end == v) { *ok = 0; return 1; }
    *out = d;
    return 1;
}

int rsvdt_opt_get_int(const char *key, PetscInt *out, int *ok)
{
    const char *v = opt_lookup(key);
    *ok = 1;
    if (!v) return 0;
    char *end;
    long long x = strtoll(v, &end, 10);
    if (end == v) {
        /* PETSc also accepts "1e3"-style integers */
        double d = strtod(v, &end);
        if (end == v || d != floor(d)) { *ok = 0; return 1; }
        x = (long long)d;
    }
    *out = (PetscInt)x;
    return 1;
}

/* PetscOptionsGetBool semantics: true/false, yes/no, on/off, 1/0; a bare key is true. */
int rsvdt_opt_get_bool(const char *key, PetscBool *out, int *ok)
{
    const char *v = opt_lookup(key);
    *ok = 1;
    if (!v) return 0;
    if (!*v || !strcasecmp(v, "true") || !strcasecmp(v, "yes") || !strcasecmp(v, "on") || !strcmp(v, "1")) {
        *out = PETSC_TRUE; return 1;
    }
    if (!strcasecmp(v, "false") || !strcasecmp(v, "no") || !strcasecmp(v, "off") || !strcmp(v, "0")) {
        *out = PETSC_FALSE; return 1;
    }
    *ok = 0;
    return 1;
}

static char *trim(char *s)
{
    while (*s && isspace((unsigned char)*s)) s++;
    char *e = s + strlen(s);
    while (e > s && isspace((unsigned char)e[-1])) *--e = 0;
    return s;
}

/* The flat "Key: value  # comment" YAML that variables.yaml uses (the keys
 * PetscOptionsInsertFileYAML turns into -Key options, SourceCode/ReadUserInput.c:22). */
PetscErrorCode rsvdt_yaml_insert_file(const char *path)
{
    FILE *f = fopen(path, "r");
    if (!f) return rsvdt_err(RSVDT_ERR_FILE, "ReadUserInput", "cannot open inputs file '%s'", path);
    char line[8192];
    int lineno = 0;
    while (fgets(line, sizeof(line), f)) {
        lineno++;
        char *s = line;
        /* strip a comment: '#' at line start or preceded by whitespace, outside quotes */
        int inq = 0; char q = 0;
        for (char *p = s; *p; p++) {
            if (inq) { if (*p == q) inq = 0; continue; }
            if (*p == '"' || *p == '\'') { inq = 1; q = *p; continue; }
            if (*p == '#' && (p == s || isspace((unsigned char)p[-1]))) { *p = 0; break; }
        }
        s = trim(s);
        if (!*s || !strcmp(s, "---") || !strcmp(s, "...")) continue;
        char *colon = strchr(s, ':');
        if (!colon) {
            fclose(f);
            return rsvdt_err(RSVDT_ERR_USER, "ReadUserInput", "%s:%d: expected 'Key: value'", path, lineno);
        }
        *colon = 0;
        char *key = trim(s), *val = trim(colon + 1);
        size_t lv = strlen(val);
        if (lv >= 2 && ((val[0] == '"' && val[lv - 1] == '"') || (val[0] == '\'' && val[lv - 1] == '\''))) {
            val[lv - 1] = 0; val++;
        }
        RSVDT_CHK(RSVDT_OptionsSetValue(key, val));
    }
    fclose(f);
    return 0;
}

/* ------------------------------------------------------------------ runtime */
static void default_options(RSVDT_Options *o)
{
    o->precision = 64;
    o->chunk_steps = 0;
    o->spmm_lanes = 0;
    o->xcd_remap = 1;
    o->use_graph = 0;
    o->autotune = 1;
    o->verbose = 0;
    o->overlap_forcing = 1;
    o->rk4_form = 0;
    o->fuse_small = 1;
    o->spmm_variant = -1;
    o->tr_group = 0;
}

PetscErrorCode RSVDT_Initialize(int *argc, char ***argv)
{
    if (!g_rsvdt.initialized) {
        memset(&g_rsvdt, 0, sizeof(g_rsvdt));
        default_options(&g_rsvdt.opt);
        g_rsvdt.device = -1;
    }
    if (argc && argv && *argv) {
        for (int i = 1; i < *argc; i++) {
            const char *a = (*argv)[i];
            if (a[0] == '-' && a[1] && !isdigit((unsigned char)a[1]) && a[1] != '.') {
                const char *val = "";
                if (i + 1 < *argc) {
                    const char *b = (*argv)[i + 1];
                    if (!(b[0] == '-' && b[1] && !isdigit((unsigned char)b[1]) && b[1] != '.')) { val = b; i++; }
                }
                RSVDT_CHK(RSVDT_OptionsSetValue(a, val));
            }
        }
    }
    g_rsvdt.initialized = 1;
    return 0;
}

PetscErrorCode rsvdt_ensure_device(void)
{
    if (!g_rsvdt.initialized) RSVDT_CHK(RSVDT_Initialize(NULL, NULL));
    if (g_rsvdt.device >= 0) return 0;
    int dev = 0;
    const char *e = getenv("RSVDT_DEVICE");
    if (e) dev = atoi(e);
    int rc = rsvdt_hip_init(dev);
    if (rc) return rsvdt_err(rc, "RSVDT_Initialize", "%s", rsvdt_hip_last_error());
    g_rsvdt.device = dev;
    rc = rsvdt_hip_stream_create(&g_rsvdt.stream);
    if (!rc) rc = rsvdt_hip_stream_create(&g_rsvdt.stream_frc);
    if (rc) return rsvdt_err(rc, "RSVDT_Initialize", "%s", rsvdt_hip_last_error());
    return 0;
}

PetscErrorCode RSVDT_Finalize(void)
{
    if (g_rsvdt.device >= 0) rsvdt_action_release();
    rsvdt_scratch_free_all();
    if (g_rsvdt.device >= 0) {
        rsvdt_hip_release_cached();
        rsvdt_hip_stream_destroy(g_rsvdt.stream);
        rsvdt_hip_stream_destroy(g_rsvdt.stream_frc);
        g_rsvdt.stream = g_rsvdt.stream_frc = NULL;
    }
    free(g_rsvdt.sketch_host);
    g_rsvdt.sketch_host = NULL;
    RSVDT_OptionsClear();
    g_rsvdt.initialized = 0;
    g_rsvdt.device = -1;
    return 0;
}

PetscErrorCode RSVDT_GetOptions(RSVDT_Options *o)
{
    if (!g_rsvdt.initialized) RSVDT_CHK(RSVDT_Initialize(NULL, NULL));
    *o = g_rsvdt.opt;
    return 0;
}

PetscErrorCode RSVDT_SetOptions(const RSVDT_Options *o)
{
    if (!g_rsvdt.initialized) RSVDT_CHK(RSVDT_Initialize(NULL, NULL));
    if (o->precision != 64 && o->precision != 32)
        return rsvdt_err(RSVDT_ERR_USER, "RSVDT_SetOptions", "precision must be 64 or 32");
    g_rsvdt.opt = *o;
    return 0;
}

PetscErrorCode RSVDT_GetActionStats(RSVDT_ActionStats *s)
{
    *s = g_rsvdt.stats;
    return 0;
}

PetscErrorCode RSVDT_SetSketchColumns(PetscInt N, PetscInt ncols, PetscInt Nw_eff, const double *host)
{
    free(g_rsvdt.sketch_host);
    g_rsvdt.sketch_host = NULL;
    g_rsvdt.sketch_dev = NULL;
    g_rsvdt.sketch_cols = 0;
    if (!host || ncols <= 0) return 0;
    size_t n = (size_t)N * (size_t)ncols * (size_t)Nw_eff * 2;
    g_rsvdt.sketch_host = malloc(n * sizeof(double));
    if (!g_rsvdt.sketch_host) return rsvdt_err(RSVDT_ERR_MEM, "RSVDT_SetSketchColumns", "out of memory");
    memcpy(g_rsvdt.sketch_host, host, n * sizeof(double));
    g_rsvdt.sketch_N = N; g_rsvdt.sketch_cols = ncols; g_rsvdt.sketch_Nw_eff = Nw_eff;
    return 0;
}

PetscErrorCode RSVDT_SetSketchColumnsDevice(PetscInt N, PetscInt ncols, PetscInt Nw_eff, const float *dev, PetscInt ld)
{
    RSVDT_CHK(RSVDT_SetSketchColumns(0, 0, 0, NULL));
    if (!dev || ncols <= 0) return 0;
    if (N <= 0 || Nw_eff <= 0 || ld < 2 * ncols)
        return rsvdt_err(RSVDT_ERR_USER, "RSVDT_SetSketchColumnsDevice", "need N, Nw_eff > 0 and ld >= 2 ncols");
    g_rsvdt.sketch_dev = dev;
    g_rsvdt.sketch_N = N; g_rsvdt.sketch_cols = ncols; g_rsvdt.sketch_Nw_eff = Nw_eff; g_rsvdt.sketch_ld = ld;
    return 0;
}

/* ------------------------------------------------------------------ scratch */
#define NSCRATCH RSVDT_NSCRATCH
static void *g_scr[NSCRATCH];
static size_t g_scr_sz[NSCRATCH];

void *rsvdt_scratch(int slot, size_t bytes)
{
    if (slot < 0 || slot >= NSCRATCH) return NULL;
    if (g_scr_sz[slot] >= bytes && g_scr[slot]) return g_scr[slot];
    if (g_scr[slot]) { rsvdt_hip_sync(); rsvdt_hip_free(g_scr[slot]); }
    g_scr[slot] = NULL; g_scr_sz[slot] = 0;
    if (rsvdt_hip_malloc(&g_scr[slot], bytes)) return NULL;
    g_scr_sz[slot] = bytes;
    return g_scr[slot];
}

void rsvdt_scratch_release(int slot)
{
    if (slot < 0 || slot >= NSCRATCH || !g_scr[slot]) return;
    rsvdt_hip_sync();
    rsvdt_hip_free(g_scr[slot]);
    g_scr[slot] = NULL; g_scr_sz[slot] = 0;
}

size_t rsvdt_scratch_size(int slot)
{
    return slot >= 0 && slot < NSCRATCH && g_scr[slot] ? g_scr_sz[slot] : 0;
}

void rsvdt_scratch_free_all(void)
{
    if (g_rsvdt.device >= 0) rsvdt_hip_sync();
    for (int i = 0; i < NSCRATCH; i++) {
        if (g_scr[i]) rsvdt_hip_free(g_scr[i]);
        g_scr[i] = NULL; g_scr_sz[i] = 0;
    }
}

/* ------------------------------------------------------------------ Mat */
PetscErrorCode rsvdt_mat_alloc_dense(PetscInt rows, PetscInt cols, PetscInt kb, int order, Mat *out)
{
    RSVDT_CHK(rsvdt_ensure_device());
    if (rows < 0 || cols < 0 || kb <= 0 || (cols % kb) != 0)
        return rsvdt_err(RSVDT_ERR_USER, "MatCreate", "bad dense shape %lld x %lld (block %lld)",
                         (long long)rows, (long long)cols, (long long)kb);
    Mat m = calloc(1, sizeof(*m));
    if (!m) return rsvdt_err(RSVDT_ERR_MEM, "MatCreate", "out of memory");
    m->kind = RSVDT_MAT_DENSE;
    m->rows = rows; m->cols = cols; m->kb = kb; m->order = order;
    size_t bytes = (size_t)rows * (size_t)cols * 16;
    int rc = rsvdt_hip_malloc((void **)&m->d, bytes);
    if (rc) {
        free(m);
        return rsvdt_err(rc, "MatCreate", "cannot allocate %.3f GB of HBM: %s", bytes / 1e9, rsvdt_hip_last_error());
    }
    rc = rsvdt_hip_zero(m->d, bytes, g_rsvdt.stream);
    if (rc) return rsvdt_err(rc, "MatCreate", "%s", rsvdt_hip_last_error());
    *out = m;
    return 0;
}

/* host column-major (rows x cols) <-> blocked device layout */
static void host_to_blocked(const Mat m, const double *cm, double *blk)
{
    PetscInt nb = m->cols / m->kb;
    for (PetscInt c = 0; c < m->cols; c++) {
        PetscInt b, j;
        if (m->order == 0) { b = c / m->kb; j = c % m->kb; }
        else { j = c / nb; b = c % nb; }
        double *dst = blk + 2 * ((size_t)b * m->rows * m->kb + j);
        const double *src = cm + 2 * (size_t)c * m->rows;
        for (PetscInt r = 0; r < m->rows; r++) {
            dst[2 * (size_t)r * m->kb] = src[2 * r];
            dst[2 * (size_t)r * m->kb + 1] = src[2 * r + 1];
        }
    }
}
static void blocked_to_host(const Mat m, const double *blk, double *cm)
{
    PetscInt nb = m->cols / m->kb;
    for (PetscInt c = 0; c < m->cols; c++) {
        PetscInt b, j;
        if (m->order == 0) { b = c / m->kb; j = c % m->kb; }
        else { j = c / nb; b = c % nb; }
        const double *src = blk + 2 * ((size_t)b * m->rows * m->kb + j);
        double *dst = cm + 2 * (size_t)c * m->rows;
        for (PetscInt r = 0; r < m->rows; r++) {
            dst[2 * r] = src[2 * (size_t)r * m->kb];
            dst[2 * r + 1] = src[2 * (size_t)r * m->kb + 1];
        }
    }
}

PetscErrorCode RSVDT_MatCreateDenseBlocked(PetscInt rows, PetscInt cols, PetscInt kb, const double *host, Mat *out)
{
    Mat m;
    RSVDT_CHK(rsvdt_mat_alloc_dense(rows, cols, kb, 0, &m));
    if (host) {
        size_t n = (size_t)rows * (size_t)cols * 2;
        double *tmp = malloc(n * sizeof(double));
        if (!tmp) return rsvdt_err(RSVDT_ERR_MEM, "MatCreateDense", "out of memory");
        host_to_blocked(m, host, tmp);
        int rc = rsvdt_hip_memcpy_h2d(m->d, tmp, n * sizeof(double));
        free(tmp);
        if (rc) return rsvdt_err(rc, "MatCreateDense", "%s", rsvdt_hip_last_error());
    }
    *out = m;
    return 0;
}

PetscErrorCode RSVDT_MatCreateDense(PetscInt rows, PetscInt cols, const double *host, Mat *out)
{
    return RSVDT_MatCreateDenseBlocked(rows, cols, cols > 0 ? 1 : 1, host, out);
}

PetscErrorCode RSVDT_MatGetDense(Mat m, double *host)
{
    if (!m || m->kind != RSVDT_MAT_DENSE) return rsvdt_err(RSVDT_ERR_USER, "MatGetDense", "not a dense Mat");
    size_t n = (size_t)m->rows * (size_t)m->cols * 2;
    double *tmp = malloc(n * sizeof(double) + 16);
    if (!tmp) return rsvdt_err(RSVDT_ERR_MEM, "MatGetDense", "out of memory");
    int rc = rsvdt_hip_sync();
    if (!rc) rc = rsvdt_hip_memcpy_d2h(tmp, m->d, n * sizeof(double));
    if (rc) { free(tmp); return rsvdt_err(rc, "MatGetDense", "%s", rsvdt_hip_last_error()); }
    blocked_to_host(m, tmp, host);
    free(tmp);
    return 0;
}

PetscErrorCode RSVDT_MatGetSize(Mat m, PetscInt *rows, PetscInt *cols)
{
    if (!m) return rsvdt_err(RSVDT_ERR_USER, "MatGetSize", "null Mat");
    if (rows) *rows = m->rows;
    if (cols) *cols = m->cols;
    return 0;
}

PetscErrorCode RSVDT_MatDestroy(Mat *pm)
{
    if (!pm || !*pm) return 0;
    Mat m = *pm;
    if (g_rsvdt.device >= 0) rsvdt_hip_sync();
    if (g_rsvdt.device >= 0 || m->d || m->d_rowptr) rsvdt_hip_free(m->d);
    rsvdt_hip_free(m->d_rowptr); rsvdt_hip_free(m->d_col);
    rsvdt_hip_free(m->d_rowptrH); rsvdt_hip_free(m->d_colH);
    rsvdt_hip_free(m->d_val64);
    rsvdt_hip_free(m->d_valH64);
    rsvdt_hip_free(m->d_valz); rsvdt_hip_free(m->d_valzH);
    rsvdt_hip_free(m->d_gptr); rsvdt_hip_free(m->d_gword); rsvdt_hip_free(m->d_gval);
    rsvdt_hip_free(m->d_gptrH); rsvdt_hip_free(m->d_gwordH); rsvdt_hip_free(m->d_gvalH);
    for (int d = 0; d < 2; d++)
        for (int i = 0; i < RSVDT_NTILE; i++) {
            rsvdt_hip_free(m->tiled[d][i].d_tptr);
            rsvdt_hip_free(m->tiled[d][i].d_tcol);
            rsvdt_hip_free(m->tiled[d][i].d_lidx);
        }
    free(m->h_rowptrH); free(m->h_colH);
    free(m->h_rowptr); free(m->h_col); free(m->h_val);
    free(m);
    *pm = NULL;
    return 0;
}

PetscErrorCode rsvdt_mat_dup(Mat src, Mat *out)
{
    if (src->kind != RSVDT_MAT_DENSE) return rsvdt_err(RSVDT_ERR_USER, "MatDuplicate", "dense only");
    Mat m;
    RSVDT_CHK(rsvdt_mat_alloc_dense(src->rows, src->cols, src->kb, src->order, &m));
    RSVDT_HIPCHK("MatDuplicate", rsvdt_hip_memcpy_d2d(m->d, src->d, (size_t)src->rows * src->cols * 16, g_rsvdt.stream));
    *out = m;
    return 0;
}

/* Physically relayout a dense Mat to (kb, order), keeping its logical content. */
PetscErrorCode rsvdt_mat_to_layout(Mat m, PetscInt kb, int order)
{
    if (m->kind != RSVDT_MAT_DENSE) return rsvdt_err(RSVDT_ERR_USER, "MatRelayout", "dense only");
    if (m->kb == kb && (m->order == order || m->cols == kb)) { m->order = order; return 0; }
    if (m->cols % kb) return rsvdt_err(RSVDT_ERR_USER, "MatRelayout", "%lld columns not divisible by %lld",
                                       (long long)m->cols, (long long)kb);
    double *nd;
    size_t bytes = (size_t)m->rows * m->cols * 16;
    RSVDT_HIPCHK("MatRelayout", rsvdt_hip_malloc((void **)&nd, bytes));
    PetscInt nb_old = m->cols / m->kb, nb_new = m->cols / kb;
    /* one strided copy per logical column */
    for (PetscInt c = 0; c < m->cols; c++) {
        PetscInt bo, jo, bn, jn;
        if (m->order == 0) { bo = c / m->kb; jo = c % m->kb; } else { jo = c / nb_old; bo = c % nb_old; }
        if (order == 0) { bn = c / kb; jn = c % kb; } else { jn = c / nb_new; bn = c % nb_new; }
        RSVDT_HIPCHK("MatRelayout",
                     rsvdt_hip_zcopy3(1, m->rows, 1, m->d + 2 * ((size_t)bo * m->rows * m->kb + jo), 0, m->kb, 0,
                                      nd + 2 * ((size_t)bn * m->rows * kb + jn), 0, kb, 0, g_rsvdt.stream));
    }
    rsvdt_hip_sync();
    rsvdt_hip_free(m->d);
    m->d = nd; m->kb = kb; m->order = order;
    return 0;
}

/* ---- AIJ ---- */
PetscErrorCode rsvdt_aij_real_check(Mat m)
{
    double im2 = 0;
    for (PetscInt e = 0; e < m->nnz; e++) im2 += m->h_val[2 * e + 1] * m->h_val[2 * e + 1];
    m->is_real = sqrt(im2) < 1e-12;
    return 0;
}

static PetscErrorCode upload_csr(PetscInt rows, const int64_t *rp, const int64_t *ci, const double *v, PetscInt nnz,
                                 int is_real, int32_t **d_rp, int32_t **d_ci, void **d64)
{
    if (nnz > 0x7fffffffLL || rows > 0x7fffffffLL)
        return rsvdt_err(RSVDT_ERR_USER, "MatLoad", "operator too large for 32-bit device indices");
    int32_t *hrp = malloc(sizeof(int32_t) * (size_t)(rows + 1));
    int32_t *hci = malloc(sizeof(int32_t) * (size_t)(nnz ? nnz : 1));
    size_t nv = (size_t)(nnz ? nnz : 1) * (is_real ? 1 : 2);
    double *h64 = malloc(sizeof(double) * nv);
    if (!hrp || !hci || !h64) { free(hrp); free(hci); free(h64); return rsvdt_err(RSVDT_ERR_MEM, "MatLoad", "out of memory"); }
    for (PetscInt r = 0; r <= rows; r++) hrp[r] = (int32_t)rp[r];
    for (PetscInt e = 0; e < nnz; e++) {
        hci[e] = (int32_t)ci[e];
        if (is_real) h64[e] = v[2 * e];
        else { h64[2 * e] = v[2 * e]; h64[2 * e + 1] = v[2 * e + 1]; }
    }
    int rc = rsvdt_hip_malloc((void **)d_rp, sizeof(int32_t) * (size_t)(rows + 1));
    if (!rc) rc = rsvdt_hip_malloc((void **)d_ci, sizeof(int32_t) * (size_t)(nnz ? nnz : 1));
    if (!rc) rc = rsvdt_hip_malloc(d64, sizeof(double) * nv);
    if (!rc) rc = rsvdt_hip_memcpy_h2d(*d_rp, hrp, sizeof(int32_t) * (size_t)(rows + 1));
    if (!rc) rc = rsvdt_hip_memcpy_h2d(*d_ci, hci, sizeof(int32_t) * (size_t)nnz);
    if (!rc) rc = rsvdt_hip_memcpy_h2d(*d64, h64, sizeof(double) * nv);
    free(hrp); free(hci); free(h64);
    if (rc) return rsvdt_err(rc, "MatLoad", "%s", rsvdt_hip_last_error());
    return 0;
}

/* ---- row-grouped form (for the grouped-gather stage kernel) ----------- */
typedef struct { int64_t col; int r; int64_t e; } gent;
static int gent_cmp(const void *a, const void *b)
{
    const gent *x = a, *y = b;
    if (x->col != y->col) return x->col < y->col ? -1 : 1;
    return x->r - y->r;
}

/* Distinct columns of the groups of G rows among the first `rows` rows. */
static int64_t group_unique(PetscInt rows, const int64_t *rp, const int64_t *ci, int G, gent *buf)
{
    int64_t u = 0;
    for (PetscInt r0 = 0; r0 < rows; r0 += G) {
        int n = 0;
        for (int r = 0; r < G && r0 + r < rows; r++)
            for (int64_t e = rp[r0 + r]; e < rp[r0 + r + 1]; e++) buf[n++] = (gent){ci[e], r, e};
        qsort(buf, (size_t)n, sizeof(gent), gent_cmp);
        for (int i = 0; i < n; i++) u += (i == 0 || buf[i].col != buf[i - 1].col);
    }
    return u;
}

/* Build and upload the row-grouped form of a CSR matrix: the group size in
 * {2, 3, 4, 5, 8} with the fewest distinct columns per row on a sample of
 * the rows, kept only when it saves at least 15 % of the gathers. */
static PetscErrorCode upload_grouped(PetscInt rows, PetscInt cols, const int64_t *rp, const int64_t *ci, const double *v,
                                     int is_real, int *G_out, int32_t **d_gptr, uint32_t **d_gword, void **d_gval)
{
    *G_out = 0;
    const PetscInt nnz = rp[rows];
    if (rows < 64 || cols >= (1LL << 24) || nnz > 0x7fffffffLL) return 0;
    int64_t maxrow = 0;
    for (PetscInt r = 0; r < rows; r++) if (rp[r + 1] - rp[r] > maxrow) maxrow = rp[r + 1] - rp[r];
    if (maxrow > 64) return 0;
    gent *buf = malloc(sizeof(gent) * (size_t)(8 * maxrow + 1));
    if (!buf) return rsvdt_err(RSVDT_ERR_MEM, "MatLoad", "out of memory");
    const int cand[5] = {2, 3, 4, 5, 8};
    const PetscInt ns = rows < 65520 ? rows : 65520;    /* sample: a multiple of 2, 3, 4, 5, 8 */
    const double base = (double)(rp[ns] - rp[0]);
    int best = 0;
    double bu = base;
    for (int c = 0; c < 5; c++) {
        double u = (double)group_unique(ns, rp, ci, cand[c], buf);
        if (u < bu) { bu = u; best = cand[c]; }
    }
    if (!best || bu > 0.85 * base) { free(buf); return 0; }
    const int G = best;
    const int64_t ng = (rows + G - 1) / G;
    int32_t *gptr = malloc(sizeof(int32_t) * (size_t)(ng + 1));
    uint32_t *gword = malloc(sizeof(uint32_t) * (size_t)(nnz ? nnz : 1));
    const size_t nv = (size_t)(nnz ? nnz : 1) * (is_real ? 1 : 2);
    double *gval = malloc(sizeof(double) * nv);
    if (!gptr || !gword || !gval) { free(buf); free(gptr); free(gword); free(gval); return rsvdt_err(RSVDT_ERR_MEM, "MatLoad", "out of memory"); }
    int64_t q = 0;
    for (int64_t g = 0; g < ng; g++) {
        const PetscInt r0 = g * G;
        gptr[g] = (int32_t)q;
        int n = 0;
        for (int r = 0; r < G && r0 + r < rows; r++)
            for (int64_t e = rp[r0 + r]; e < rp[r0 + r + 1]; e++) buf[n++] = (gent){ci[e], r, e};
        qsort(buf, (size_t)n, sizeof(gent), gent_cmp);
        int64_t vo = rp[r0];
        for (int i = 0; i < n; i++) {
            if (i == 0 || buf[i].col != buf[i - 1].col) gword[q++] = ((uint32_t)buf[i].col << 8);
            gword[q - 1] |= 1u << buf[i].r;
            if (is_real) gval[vo] = v[2 * buf[i].e];
            else { gval[2 * vo] = v[2 * buf[i].e]; gval[2 * vo + 1] = v[2 * buf[i].e + 1]; }
            vo++;
        }
    }
    gptr[ng] = (int32_t)q;
    free(buf);
    int rc = rsvdt_hip_malloc((void **)d_gptr, sizeof(int32_t) * (size_t)(ng + 1));
    if (!rc) rc = rsvdt_hip_malloc((void **)d_gword, sizeof(uint32_t) * (size_t)(q ? q : 1));
    if (!rc) rc = rsvdt_hip_malloc(d_gval, sizeof(double) * nv);
    if (!rc) rc = rsvdt_hip_memcpy_h2d(*d_gptr, gptr, sizeof(int32_t) * (size_t)(ng + 1));
    if (!rc && q) rc = rsvdt_hip_memcpy_h2d(*d_gword, gword, sizeof(uint32_t) * (size_t)q);
    if (!rc) rc = rsvdt_hip_memcpy_h2d(*d_gval, gval, sizeof(double) * nv);
    free(gptr); free(gword); free(gval);
    if (rc) return rsvdt_err(rc, "MatLoad", "%s", rsvdt_hip_last_error());
    *G_out = G;
    return 0;
}

/* ---- tiled form (for the LDS-staged stage kernel) --------------------- */
const int rsvdt_tile_rows[RSVDT_NTILE] = {8, 16, 32, 48, 64};

typedef struct { int64_t col, e; } tent;
static int tent_cmp(const void *a, const void *b)
{
    const tent *x = a, *y = b;
    if (x->col != y->col) return x->col < y->col ? -1 : 1;
    return x->e < y->e ? -1 : (x->e > y->e);
}

/* Tiles of TR consecutive rows: the sorted distinct columns of each tile and
 * every nonzero's position among them. */
static PetscErrorCode build_tiled(PetscInt rows, const int64_t *rp, const int64_t *ci, int TR, rsvdt_tiled *t)
{
    const int64_t nt = (rows + TR - 1) / TR, nnz = rp[rows];
    if (nnz > 0x7fffffffLL || rows > 0x7fffffffLL) { t->rows = -1; return 0; }
    int64_t emax = 0;
    for (int64_t q = 0; q < nt; q++) {
        const int64_t r0 = q * TR, r1 = r0 + TR < rows ? r0 + TR : rows;
        if (rp[r1] - rp[r0] > emax) emax = rp[r1] - rp[r0];
    }
    int32_t *tptr = malloc(sizeof(int32_t) * (size_t)(nt + 1));
    int32_t *tcol = malloc(sizeof(int32_t) * (size_t)(nnz ? nnz : 1));
    uint16_t *lidx = malloc(sizeof(uint16_t) * (size_t)(nnz ? nnz : 1));
    tent *buf = malloc(sizeof(tent) * (size_t)(emax ? emax : 1));
    if (!tptr || !tcol || !lidx || !buf) {
        free(tptr); free(tcol); free(lidx); free(buf);
        return rsvdt_err(RSVDT_ERR_MEM, "MatAIJ", "out of host memory (tiled form)");
    }
    int64_t h = 0, hmax = 0;
    int ok = 1;
    for (int64_t q = 0; q < nt && ok; q++) {
        const int64_t r0 = q * TR, r1 = r0 + TR < rows ? r0 + TR : rows;
        const int64_t e0 = rp[r0], n = rp[r1] - e0;
        for (int64_t i = 0; i < n; i++) buf[i] = (tent){ci[e0 + i], e0 + i};
        qsort(buf, (size_t)n, sizeof(tent), tent_cmp);
        tptr[q] = (int32_t)h;
        for (int64_t i = 0; i < n; i++) {
            if (i == 0 || buf[i].col != buf[i - 1].col) tcol[h++] = (int32_t)buf[i].col;
            const int64_t slot = h - 1 - tptr[q];
            if (slot > 65535) { ok = 0; break; }
            lidx[buf[i].e] = (uint16_t)slot;
        }
        if (h - tptr[q] > hmax) hmax = h - tptr[q];
    }
    tptr[nt] = (int32_t)h;
    free(buf);
    if (!ok) {
        free(tptr); free(tcol); free(lidx);
        t->rows = -1;
        return 0;
    }
    int rc = rsvdt_hip_malloc((void **)&t->d_tptr, sizeof(int32_t) * (size_t)(nt + 1));
    if (!rc) rc = rsvdt_hip_malloc((void **)&t->d_tcol, sizeof(int32_t) * (size_t)(h ? h : 1));
    if (!rc) rc = rsvdt_hip_malloc((void **)&t->d_lidx, sizeof(uint16_t) * (size_t)(nnz ? nnz : 1));
    if (!rc) rc = rsvdt_hip_memcpy_h2d(t->d_tptr, tptr, sizeof(int32_t) * (size_t)(nt + 1));
    if (!rc && h) rc = rsvdt_hip_memcpy_h2d(t->d_tcol, tcol, sizeof(int32_t) * (size_t)h);
    if (!rc && nnz) rc = rsvdt_hip_memcpy_h2d(t->d_lidx, lidx, sizeof(uint16_t) * (size_t)nnz);
    free(tptr); free(tcol); free(lidx);
    if (rc) {
        rsvdt_hip_free(t->d_tptr); rsvdt_hip_free(t->d_tcol); rsvdt_hip_free(t->d_lidx);
        t->d_tptr = t->d_tcol = NULL; t->d_lidx = NULL;
        return rsvdt_err(rc, "MatAIJ", "%s", rsvdt_hip_last_error());
    }
    t->rows = TR;
    t->hmax = (int)hmax;
    t->emax = (int)emax;
    return 0;
}

PetscErrorCode rsvdt_tiled_get(Mat m, int dir, int tile_rows, const rsvdt_tiled **out)
{
    *out = NULL;
    RSVDT_CHK(rsvdt_aij_upload(m));
    int ix = -1;
    for (int i = 0; i < RSVDT_NTILE; i++) if (rsvdt_tile_rows[i] == tile_rows) ix = i;
    if (ix < 0) return 0;
    const int d = dir ? 0 : 1;
    rsvdt_tiled *t = &m->tiled[d][ix];
    if (t->rows == 0) {
        if (dir) RSVDT_CHK(build_tiled(m->rows, m->h_rowptr, m->h_col, tile_rows, t));
        else RSVDT_CHK(build_tiled(m->cols, m->h_rowptrH, m->h_colH, tile_rows, t));
    }
    if (t->rows > 0) *out = t;
    return 0;
}

/* Upload A and A^H (built on the host by a counting transpose).  Called
 * lazily on first device use, so that loading and validating matrices needs
 * no GPU. */
PetscErrorCode rsvdt_aij_upload(Mat m)
{
    if (m->kind != RSVDT_MAT_AIJ) return rsvdt_err(RSVDT_ERR_USER, "MatAIJ", "not an AIJ Mat");
    if (m->d_rowptr) return 0;
    RSVDT_CHK(rsvdt_ensure_device());
    RSVDT_CHK(rsvdt_aij_real_check(m));
    RSVDT_CHK(upload_csr(m->rows, m->h_rowptr, m->h_col, m->h_val, m->nnz, m->is_real, &m->d_rowptr, &m->d_col,
                         &m->d_val64));
    RSVDT_CHK(upload_grouped(m->rows, m->cols, m->h_rowptr, m->h_col, m->h_val, m->is_real, &m->group, &m->d_gptr,
                             &m->d_gword, &m->d_gval));
    m->maxrow = 0;
    for (PetscInt r = 0; r < m->rows; r++)
        if (m->h_rowptr[r + 1] - m->h_rowptr[r] > m->maxrow) m->maxrow = m->h_rowptr[r + 1] - m->h_rowptr[r];
    PetscInt nr = m->cols, nnz = m->nnz;
    int64_t *rp = calloc((size_t)nr + 1, sizeof(int64_t));
    int64_t *ci = malloc(sizeof(int64_t) * (size_t)(nnz ? nnz : 1));
    double *v = malloc(sizeof(double) * 2 * (size_t)(nnz ? nnz : 1));
    int64_t *fill = malloc(sizeof(int64_t) * (size_t)(nr ? nr : 1));
    if (!rp || !ci || !v || !fill) return rsvdt_err(RSVDT_ERR_MEM, "MatLoad", "out of memory");
    for (PetscInt e = 0; e < nnz; e++) rp[m->h_col[e] + 1]++;
    for (PetscInt r = 0; r < nr; r++) rp[r + 1] += rp[r];
    memcpy(fill, rp, sizeof(int64_t) * (size_t)nr);
    for (PetscInt r = 0; r < m->rows; r++)
        for (int64_t e = m->h_rowptr[r]; e < m->h_rowptr[r + 1]; e++) {
            int64_t c = m->h_col[e];
            int64_t p = fill[c]++;
            ci[p] = r;
            v[2 * p] = m->h_val[2 * e];
            v[2 * p + 1] = -m->h_val[2 * e + 1];
        }
    PetscErrorCode ierr = upload_csr(nr, rp, ci, v, nnz, m->is_real, &m->d_rowptrH, &m->d_colH, &m->d_valH64);
    if (!ierr) ierr = upload_grouped(nr, m->rows, rp, ci, v, m->is_real, &m->groupH, &m->d_gptrH, &m->d_gwordH,
                                     &m->d_gvalH);
    m->maxrowH = 0;
    for (PetscInt r = 0; r < nr; r++)
        if (rp[r + 1] - rp[r] > m->maxrowH) m->maxrowH = rp[r + 1] - rp[r];
    free(m->h_rowptrH); free(m->h_colH);
    m->h_rowptrH = rp;      /* kept for the tiled form of A^H */
    m->h_colH = ci;
    free(v); free(fill);
    memset(m->tuned_ok, 0, sizeof(m->tuned_ok));
    memset(m->mt_ok, 0, sizeof(m->mt_ok));
    return ierr;
}

PetscErrorCode RSVDT_MatCreateAIJ(PetscInt rows, PetscInt cols, const int64_t *rowptr, const int64_t *colidx,
                                  const double *vals, Mat *out)
{
    Mat m = calloc(1, sizeof(*m));
    if (!m) return rsvdt_err(RSVDT_ERR_MEM, "MatCreateAIJ", "out of memory");
    m->kind = RSVDT_MAT_AIJ;
    m->rows = rows; m->cols = cols;
    m->nnz = rowptr[rows];
    m->h_rowptr = malloc(sizeof(int64_t) * (size_t)(rows + 1));
    m->h_col = malloc(sizeof(int64_t) * (size_t)(m->nnz ? m->nnz : 1));
    m->h_val = malloc(sizeof(double) * 2 * (size_t)(m->nnz ? m->nnz : 1));
    if (!m->h_rowptr || !m->h_col || !m->h_val) return rsvdt_err(RSVDT_ERR_MEM, "MatCreateAIJ", "out of memory");
    memcpy(m->h_rowptr, rowptr, sizeof(int64_t) * (size_t)(rows + 1));
    for (PetscInt e = 0; e < m->nnz; e++) {
        if (colidx[e] < 0 || colidx[e] >= cols) {
            RSVDT_MatDestroy(&m);
            return rsvdt_err(RSVDT_ERR_USER, "MatCreateAIJ", "column index %lld out of range", (long long)colidx[e]);
        }
        m->h_col[e] = colidx[e];
    }
    memcpy(m->h_val, vals, sizeof(double) * 2 * (size_t)m->nnz);
    RSVDT_CHK(rsvdt_aij_real_check(m));
    *out = m;
    return 0;
}

PetscErrorCode RSVDT_MatGetAIJ(Mat m, PetscInt *nnz, int64_t *rowptr, int64_t *colidx, double *vals, int *is_real)
{
    if (!m || m->kind != RSVDT_MAT_AIJ) return rsvdt_err(RSVDT_ERR_USER, "MatGetAIJ", "not an AIJ Mat");
    if (nnz) *nnz = m->nnz;
    if (rowptr) memcpy(rowptr, m->h_rowptr, sizeof(int64_t) * (size_t)(m->rows + 1));
    if (colidx) memcpy(colidx, m->h_col, sizeof(int64_t) * (size_t)m->nnz);
    if (vals) memcpy(vals, m->h_val, sizeof(double) * 2 * (size_t)m->nnz);
    if (is_real) *is_real = m->is_real;
    return 0;
}

/* ------------------------------------------------------------------ PETSc binary I/O */
#define MAT_FILE_CLASSID 1211216

static int64_t be64(const unsigned char *p)
{
    uint64_t v = 0;
    for (int i = 0; i < 8; i++) v = (v << 8) | p[i];
    return (int64_t)v;
}
static int32_t be32(const unsigned char *p)
{
    uint32_t v = 0;
    for (int i = 0; i < 4; i++) v = (v << 8) | p[i];
    return (int32_t)v;
}
static double bedbl(const unsigned char *p)
{
    uint64_t v = (uint64_t)be64(p);
    double d;
    memcpy(&d, &v, 8);
    return d;
}
static void put_be64(unsigned char *p, uint64_t v)
{
    for (int i = 7; i >= 0; i--) { p[i] = (unsigned char)(v & 0xff); v >>= 8; }
}

/* Reads the first Mat object of a PETSc binary file (a trailing object, such
 * as the Vec that follows the matrix in Tutorial/Ginzburg-Landau/A_GL, is
 * ignored as MatLoad ignores it).  64- or 32-bit indices are detected from
 * the header; complex (16-byte) or real (8-byte) scalars from the file size.
 * as_sparse: keep AIJ (operators, weights); otherwise a dense Mat. */
PetscErrorCode RSVDT_MatLoad(const char *path, int as_sparse, Mat *out)
{
    FILE *f = fopen(path, "rb");
    if (!f) return rsvdt_err(RSVDT_ERR_FILE, "MatLoad", "cannot open '%s'", path);
    fseek(f, 0, SEEK_END);
    long long fsize = ftell(f);
    fseek(f, 0, SEEK_SET);
    unsigned char hb[32];
    if (fread(hb, 1, 32, f) < 16) { fclose(f); return rsvdt_err(RSVDT_ERR_FILE, "MatLoad", "'%s': truncated header", path); }
    int ib;                        /* index bytes */
    int64_t hdr[4];
    if (be64(hb) == MAT_FILE_CLASSID) {
        ib = 8;
        for (int i = 0; i < 4; i++) hdr[i] = be64(hb + 8 * i);
    } else if (be32(hb) == MAT_FILE_CLASSID) {
        ib = 4;
        for (int i = 0; i < 4; i++) hdr[i] = be32(hb + 4 * i);
    } else {
        fclose(f);
        return rsvdt_err(RSVDT_ERR_FILE, "MatLoad", "'%s' is not a PETSc binary Mat (bad classid)", path);
    }
    int64_t M = hdr[1], Ncol = hdr[2], nz = hdr[3];
    if (M < 0 || Ncol < 0 || (nz < 0 && nz != -1)) {
        fclose(f);
        return rsvdt_err(RSVDT_ERR_FILE, "MatLoad", "'%s': bad sizes in header", path);
    }
    long long off = 4LL * ib;
    fseek(f, off, SEEK_SET);
    int sb;                        /* scalar bytes */
    if (nz == -1) {
        long long need_c = off + (long long)M * Ncol * 16;
        sb = (fsize >= need_c) ? 16 : 8;
        if (sb == 8 && fsize != off + (long long)M * Ncol * 8) {
            fclose(f);
            return rsvdt_err(RSVDT_ERR_FILE, "MatLoad", "'%s': truncated dense data", path);
        }
    } else {
        long long need_c = off + (long long)(M + nz) * ib + (long long)nz * 16;
        long long need_r = off + (long long)(M + nz) * ib + (long long)nz * 8;
        /* complex scalars (the reference's build); an 8-byte-scalar file is
         * accepted only when its size matches that layout exactly, so a
         * truncated complex file is reported as truncated */
        if (fsize >= need_c) sb = 16;
        else if (fsize == need_r) sb = 8;
        else { fclose(f); return rsvdt_err(RSVDT_ERR_FILE, "MatLoad", "'%s': truncated matrix data", path); }
    }
    int64_t *rp = malloc(sizeof(int64_t) * (size_t)(M + 1));
    int64_t nnz = (nz == -1) ? M * Ncol : nz;
    int64_t *ci = malloc(sizeof(int64_t) * (size_t)(nnz ? nnz : 1));
    double *v = malloc(sizeof(double) * 2 * (size_t)(nnz ? nnz : 1));
    unsigned char *buf = malloc((size_t)((nnz > M ? nnz : M) + 1) * (size_t)(sb > ib ? sb : ib));
    if (!rp || !ci || !v || !buf) { fclose(f); return rsvdt_err(RSVDT_ERR_MEM, "MatLoad", "out of memory"); }
    PetscErrorCode ierr = 0;
    rp[0] = 0;
    if (nz == -1) {
        for (int64_t r = 0; r < M; r++) rp[r + 1] = rp[r] + Ncol;
        for (int64_t e = 0; e < nnz; e++) ci[e] = e % Ncol;
    } else {
        if (fread(buf, (size_t)ib, (size_t)M, f) != (size_t)M) ierr = 1;
        for (int64_t r = 0; r < M && !ierr; r++) {
            int64_t l = ib == 8 ? be64(buf + 8 * r) : be32(buf + 4 * r);
            if (l < 0 || l > Ncol) ierr = 2;
            rp[r + 1] = rp[r] + l;
        }
        if (!ierr && rp[M] != nz) ierr = 3;
        if (!ierr && nnz && fread(buf, (size_t)ib, (size_t)nnz, f) != (size_t)nnz) ierr = 1;
        for (int64_t e = 0; e < nnz && !ierr; e++) {
            ci[e] = ib == 8 ? be64(buf + 8 * e) : be32(buf + 4 * e);
            if (ci[e] < 0 || ci[e] >= Ncol) ierr = 4;
        }
    }
    if (!ierr && nnz && fread(buf, (size_t)sb, (size_t)nnz, f) != (size_t)nnz) ierr = 1;
    for (int64_t e = 0; e < nnz && !ierr; e++) {
        if (sb == 16) { v[2 * e] = bedbl(buf + 16 * e); v[2 * e + 1] = bedbl(buf + 16 * e + 8); }
        else { v[2 * e] = bedbl(buf + 8 * e); v[2 * e + 1] = 0; }
    }
    fclose(f);
    free(buf);
    if (ierr) {
        free(rp); free(ci); free(v);
        return rsvdt_err(RSVDT_ERR_FILE, "MatLoad", "'%s': corrupt matrix data (code %d)", path, (int)ierr);
    }
    for (int64_t e = 0; e < nnz; e++)
        if (!isfinite(v[2 * e]) || !isfinite(v[2 * e + 1])) {
            free(rp); free(ci); free(v);
            return rsvdt_err(RSVDT_ERR_USER, "MatLoad", "'%s': non-finite value at entry %lld", path, (long long)e);
        }
    if (as_sparse) {
        ierr = RSVDT_MatCreateAIJ(M, Ncol, rp, ci, v, out);
    } else {
        double *cm = calloc((size_t)M * (size_t)Ncol * 2 + 2, sizeof(double));
        if (!cm) ierr = rsvdt_err(RSVDT_ERR_MEM, "MatLoad", "out of memory");
        else {
            for (int64_t r = 0; r < M; r++)
                for (int64_t e = rp[r]; e < rp[r + 1]; e++) {
                    cm[2 * ((size_t)ci[e] * M + r)] = v[2 * e];
                    cm[2 * ((size_t)ci[e] * M + r) + 1] = v[2 * e + 1];
                }
            ierr = RSVDT_MatCreateDense(M, Ncol, cm, out);
            free(cm);
        }
    }
    free(rp); free(ci); free(v);
    return ierr;
}

/* MatView of a dense Mat: PETSc's binary layout for MATDENSE is the AIJ one
 * with every entry stored (header [classid, M, N, M*N], M row lengths N,
 * column indices, row-major values), big-endian, 64-bit indices. */
PetscErrorCode RSVDT_MatView(Mat m, const char *path)
{
    if (!m) return rsvdt_err(RSVDT_ERR_USER, "MatView", "null Mat");
    FILE *f = fopen(path, "wb");
    if (!f) return rsvdt_err(RSVDT_ERR_FILE, "MatView", "cannot create '%s'", path);
    int64_t M = m->rows, Nc = m->cols;
    unsigned char b8[8];
    PetscErrorCode ierr = 0;
    if (m->kind == RSVDT_MAT_DENSE) {
        double *cm = malloc(sizeof(double) * 2 * (size_t)(M * Nc + 1));
        if (!cm) { fclose(f); return rsvdt_err(RSVDT_ERR_MEM, "MatView", "out of memory"); }
        ierr = RSVDT_MatGetDense(m, cm);
        if (!ierr) {
            int64_t hdr[4] = {MAT_FILE_CLASSID, M, Nc, M * Nc};
            for (int i = 0; i < 4; i++) { put_be64(b8, (uint64_t)hdr[i]); fwrite(b8, 1, 8, f); }
            unsigned char *row = malloc((size_t)(Nc > 0 ? Nc : 1) * 16);
            for (int64_t r = 0; r < M; r++) { put_be64(b8, (uint64_t)Nc); fwrite(b8, 1, 8, f); }
            for (int64_t c = 0; c < Nc; c++) put_be64(row + 8 * c, (uint64_t)c);
            for (int64_t r = 0; r < M; r++) fwrite(row, 8, (size_t)Nc, f);
            for (int64_t r = 0; r < M; r++) {
                for (int64_t c = 0; c < Nc; c++) {
                    uint64_t u;
                    memcpy(&u, &cm[2 * ((size_t)c * M + r)], 8); put_be64(row + 16 * c, u);
                    memcpy(&u, &cm[2 * ((size_t)c * M + r) + 1], 8); put_be64(row + 16 * c + 8, u);
                }
                fwrite(row, 16, (size_t)Nc, f);
            }
            free(row);
        }
        free(cm);
    } else {
        int64_t hdr[4] = {MAT_FILE_CLASSID, M, Nc, m->nnz};
        for (int i = 0; i < 4; i++) { put_be64(b8, (uint64_t)hdr[i]); fwrite(b8, 1, 8, f); }
        for (int64_t r = 0; r < M; r++) { put_be64(b8, (uint64_t)(m->h_rowptr[r + 1] - m->h_rowptr[r])); fwrite(b8, 1, 8, f); }
        for (int64_t e = 0; e < m->nnz; e++) { put_be64(b8, (uint64_t)m->h_col[e]); fwrite(b8, 1, 8, f); }
        for (int64_t e = 0; e < m->nnz; e++) {
            uint64_t u;
            memcpy(&u, &m->h_val[2 * e], 8); put_be64(b8, u); fwrite(b8, 1, 8, f);
            memcpy(&u, &m->h_val[2 * e + 1], 8); put_be64(b8, u); fwrite(b8, 1, 8, f);
        }
    }
    if (fclose(f) != 0 && !ierr) ierr = rsvdt_err(RSVDT_ERR_FILE, "MatView", "write error on '%s'", path);
    return ierr;
}
