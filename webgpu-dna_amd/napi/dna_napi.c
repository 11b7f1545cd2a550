/*
 * dna_napi.c — Node-API addon over the C ABI of libdna_amd.so (include/dna_amd.h), for
 * the reference's JavaScript/TypeScript host.
 *
 * Every GPU call runs as napi_async_work on the libuv pool and returns a Promise, so HIP
 * work never blocks the JS thread (the reference keeps its main thread free the same way:
 * WebGPU is asynchronous and IRT runs in a Worker, src/chemistry/worker.ts:32-87).
 * Outputs are fresh ArrayBuffers owned by JS, like the reference's `.slice(0)` copies of
 * mapped ranges (src/gpu/dispatch.ts:122,157,175,237).  Calls on one context are
 * serialised by a per-context mutex, and each pending call holds a reference on its
 * context object, so garbage collection cannot destroy a context under queued work.  A missing device resolves ctxCreate to null, as
 * initGPU returns null (src/gpu/device.ts:12-21); any other failure rejects with the
 * library's error text.
 *
 * Exports:
 *   version() -> string
 *   lastError() -> string
 *   ctxCreate(device, tablePath?) -> context | null
 *   ctxDestroy(context)
 *   runTransport(ctx, params, energies: Float32Array|null, copyFlags) -> Promise<object>
 *   runIrt(ctx, rad: Float32Array|null, radN, nTherm, E_eV, tCkpt: Float64Array, seed) -> Promise<object>
 *   score(ctx, target, rad: Float32Array|null, radN, lcgState, wantHits) -> Promise<object>
 *   hashCells(ctx, pos4: Float32Array, cellNm, hashBits) -> Promise<object>
 *   setOption(ctx, option, value) / getOption(ctx, option) -> number   (dna_set_option)
 *   irtOhFinal(ctx, copy?) -> Promise<{n, oh: Float32Array|null}>     (dna_irt_oh_final)
 *   irtParticles(ctx) -> Promise<{n, particles: Float64Array}>         (dna_irt_particles)
 *   scoreOh(ctx, target, rad|null, radN, oh|null, nOh, lcgState, wantHits?) -> Promise<object>
 *                                                                       (dna_score_oh)
 *   irtReactions() -> Float64Array(54)                                  (dna_irt_reactions)
 */
#define NAPI_VERSION 6
#include <node_api.h>
#include <pthread.h>
#include <stdio.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/dna_amd.h"

#define CHECK(env, call) do { napi_status s_ = (call); if (s_ != napi_ok) { \
    napi_throw_error((env), NULL, "N-API call failed: " #call); return NULL; } } while (0)

typedef struct {
    dna_ctx *ctx;
    pthread_mutex_t mu;
} Handle;

static void handle_finalize(napi_env env, void *data, void *hint)
{
    (void)env; (void)hint;
    Handle *h = (Handle *)data;
    /* every pending request holds a reference on the context object, so none is running;
       the lock orders this after the last request's unlock all the same */
    pthread_mutex_lock(&h->mu);
    if (h->ctx) dna_ctx_destroy(h->ctx);
    h->ctx = NULL;
    pthread_mutex_unlock(&h->mu);
    pthread_mutex_destroy(&h->mu);
    free(h);
}

static Handle *get_handle(napi_env env, napi_value v)
{
    void *p = NULL;
    if (napi_get_value_external(env, v, &p) != napi_ok || !p || !((Handle *)p)->ctx) {
        napi_throw_type_error(env, NULL, "expected a live dna context");
        return NULL;
    }
    return (Handle *)p;
}

/* ---------------------------------------------------------------- helpers */
static double get_num(napi_env env, napi_value obj, const char *key, double dflt)
{
    napi_value v;
    bool has = false;
    if (napi_has_named_property(env, obj, key, &has) != napi_ok || !has) return dflt;
    if (napi_get_named_property(env, obj, key, &v) != napi_ok) return dflt;
    double d = dflt;
    napi_valuetype t;
    napi_typeof(env, v, &t);
    if (t == napi_number) napi_get_value_double(env, v, &d);
    else if (t == napi_boolean) { bool b; napi_get_value_bool(env, v, &b); d = b ? 1.0 : 0.0; }
    return d;
}

/* a typed array argument (or null/undefined): data pointer, element count, and a
 * reference that keeps it alive until the async work completes */
static int get_typed(napi_env env, napi_value v, napi_typedarray_type want, void **data, size_t *len,
                     napi_ref *ref)
{
    napi_valuetype t;
    napi_typeof(env, v, &t);
    *data = NULL; *len = 0; *ref = NULL;
    if (t == napi_null || t == napi_undefined) return 0;
    bool is = false;
    napi_is_typedarray(env, v, &is);
    if (!is) return -1;
    napi_typedarray_type tt;
    napi_value ab;
    size_t off;
    if (napi_get_typedarray_info(env, v, &tt, len, data, &ab, &off) != napi_ok || tt != want) return -1;
    napi_create_reference(env, v, 1, ref);
    return 0;
}

static napi_value make_ab_copy(napi_env env, const void *src, size_t bytes, napi_typedarray_type tt, size_t n)
{
    void *dst = NULL;
    napi_value ab, arr;
    if (napi_create_arraybuffer(env, bytes, &dst, &ab) != napi_ok) return NULL;
    if (bytes) memcpy(dst, src, bytes);
    if (tt == (napi_typedarray_type)-1) return ab;
    if (napi_create_typedarray(env, tt, n, ab, 0, &arr) != napi_ok) return NULL;
    return arr;
}

static void set_num(napi_env env, napi_value obj, const char *k, double x)
{
    napi_value v;
    napi_create_double(env, x, &v);
    napi_set_named_property(env, obj, k, v);
}

static void set_val(napi_env env, napi_value obj, const char *k, napi_value v)
{
    if (!v) napi_get_null(env, &v);
    napi_set_named_property(env, obj, k, v);
}

/* common async request */
typedef struct Req Req;
typedef int (*exec_fn)(Req *);
typedef napi_value (*done_fn)(napi_env, Req *);
struct Req {
    napi_async_work work;
    napi_deferred deferred;
    Handle *h;
    napi_ref hold;          /* the context object: not finalized while the request is pending */
    napi_ref keep[6];
    exec_fn exec;
    done_fn done;
    int rc;
    char err[512];
    /* transport */
    dna_params p;
    const float *energies;
    uint32_t copy;
    dna_transport_out tout;
    void *res_copy, *dose_copy, *rad_copy;
    size_t res_bytes, dose_n, rad_n;
    /* irt */
    const float *rad;
    uint64_t rad_count;
    uint32_t n_therm;
    float E;
    const double *tck;
    int nck;
    uint64_t seed;
    dna_chem_out cout;
    /* score */
    dna_dna_target tgt;
    uint32_t lcg;
    int want_hits;
    dna_score_out sout;
    uint8_t *hits_copy;
    size_t hits_n;
    /* final IRT state and indirect scoring */
    const float *ohin;
    uint64_t oh_count;
    dna_indirect_out ind;
    int want_copy;
    float *oh_copy;
    double *parts_copy;
    uint64_t n_out;
    /* hash */
    const float *pos;
    uint64_t npos;
    float cell;
    uint32_t bits;
    uint32_t *key, *start, *order;
};

static void exec_cb(napi_env env, void *data)
{
    (void)env;
    Req *r = (Req *)data;
    pthread_mutex_lock(&r->h->mu);
    r->rc = r->exec(r);
    if (r->rc != DNA_OK) {
        const char *m = dna_last_error();
        strncpy(r->err, m ? m : "unknown error", sizeof(r->err) - 1);
    }
    pthread_mutex_unlock(&r->h->mu);
}

static void free_req(napi_env env, Req *r)
{
    for (int i = 0; i < 6; i++) if (r->keep[i]) napi_delete_reference(env, r->keep[i]);
    if (r->hold) napi_delete_reference(env, r->hold);
    free(r->res_copy); free(r->dose_copy); free(r->rad_copy); free(r->hits_copy);
    free(r->key); free(r->start); free(r->order);
    free(r->oh_copy); free(r->parts_copy);
    free(r);
}

static void complete_cb(napi_env env, napi_status status, void *data)
{
    Req *r = (Req *)data;
    if (status != napi_ok || r->rc != DNA_OK) {
        napi_value msg, err;
        char buf[600];
        snprintf(buf, sizeof(buf), "dna_amd error %d: %s", r->rc, r->rc ? r->err : "async work cancelled");
        napi_create_string_utf8(env, buf, NAPI_AUTO_LENGTH, &msg);
        napi_create_error(env, NULL, msg, &err);
        napi_reject_deferred(env, r->deferred, err);
    } else {
        napi_value v = r->done(env, r);
        if (!v) napi_get_null(env, &v);
        napi_resolve_deferred(env, r->deferred, v);
    }
    napi_delete_async_work(env, r->work);
    free_req(env, r);
}

static napi_value start(napi_env env, Req *r, const char *name)
{
    napi_value promise, rname;
    napi_create_promise(env, &r->deferred, &promise);
    napi_create_string_utf8(env, name, NAPI_AUTO_LENGTH, &rname);
    napi_create_async_work(env, NULL, rname, exec_cb, complete_cb, r, &r->work);
    napi_queue_async_work(env, r->work);
    return promise;
}

/* ---------------------------------------------------------------- version / errors */
static napi_value js_version(napi_env env, napi_callback_info info)
{
    (void)info;
    napi_value v;
    CHECK(env, napi_create_string_utf8(env, dna_version(), NAPI_AUTO_LENGTH, &v));
    return v;
}

static napi_value js_last_error(napi_env env, napi_callback_info info)
{
    (void)info;
    napi_value v;
    const char *m = dna_last_error();
    CHECK(env, napi_create_string_utf8(env, m ? m : "", NAPI_AUTO_LENGTH, &v));
    return v;
}

/* ---------------------------------------------------------------- context */
static napi_value js_ctx_create(napi_env env, napi_callback_info info)
{
    size_t argc = 2;
    napi_value argv[2];
    CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    int32_t dev = 0;
    if (argc > 0) napi_get_value_int32(env, argv[0], &dev);
    char path[4096];
    const char *tp = NULL;
    if (argc > 1) {
        napi_valuetype t;
        napi_typeof(env, argv[1], &t);
        if (t == napi_string) {
            size_t n;
            napi_get_value_string_utf8(env, argv[1], path, sizeof(path), &n);
            tp = path;
        }
    }
    dna_ctx *c = NULL;
    int rc = dna_ctx_create(dev, tp, &c);
    napi_value out;
    if (rc == DNA_ERR_NO_DEVICE) { napi_get_null(env, &out); return out; }
    if (rc != DNA_OK) { napi_throw_error(env, NULL, dna_last_error()); return NULL; }
    Handle *h = (Handle *)calloc(1, sizeof(Handle));
    h->ctx = c;
    pthread_mutex_init(&h->mu, NULL);
    CHECK(env, napi_create_external(env, h, handle_finalize, NULL, &out));
    return out;
}

static napi_value js_ctx_destroy(napi_env env, napi_callback_info info)
{
    size_t argc = 1;
    napi_value argv[1], und;
    CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    void *p = NULL;
    if (argc > 0 && napi_get_value_external(env, argv[0], &p) == napi_ok && p) {
        Handle *h = (Handle *)p;
        pthread_mutex_lock(&h->mu);
        if (h->ctx) dna_ctx_destroy(h->ctx);
        h->ctx = NULL;
        pthread_mutex_unlock(&h->mu);
    }
    napi_get_undefined(env, &und);
    return und;
}

/* ---------------------------------------------------------------- transport */
static int exec_transport(Req *r)
{
    int rc = dna_run_transport(r->h->ctx, &r->p, r->energies, r->copy, &r->tout);
    if (rc) return rc;
    if (r->tout.results) {
        r->res_bytes = (size_t)r->p.n * sizeof(dna_result);
        r->res_copy = malloc(r->res_bytes ? r->res_bytes : 1);
        memcpy(r->res_copy, r->tout.results, r->res_bytes);
    }
    if (r->tout.dose) {
        r->dose_n = (size_t)r->p.vc * r->p.vc * r->p.vc;
        r->dose_copy = malloc(r->dose_n * 4 + 4);
        memcpy(r->dose_copy, r->tout.dose, r->dose_n * 4);
    }
    if (r->tout.rad_buf) {
        r->rad_n = (size_t)r->tout.rad_n_stored;
        r->rad_copy = malloc(r->rad_n * 16 + 16);
        memcpy(r->rad_copy, r->tout.rad_buf, r->rad_n * 16);
    }
    return DNA_OK;
}

static napi_value done_transport(napi_env env, Req *r)
{
    napi_value o;
    napi_create_object(env, &o);
    const dna_transport_out *t = &r->tout;
    set_val(env, o, "results", r->res_copy ? make_ab_copy(env, r->res_copy, r->res_bytes, (napi_typedarray_type)-1, 0) : NULL);
    set_val(env, o, "dose", r->dose_copy ? make_ab_copy(env, r->dose_copy, r->dose_n * 4, napi_uint32_array, r->dose_n) : NULL);
    set_val(env, o, "rad", r->rad_copy ? make_ab_copy(env, r->rad_copy, r->rad_n * 16, napi_float32_array, r->rad_n * 4) : NULL);
    set_val(env, o, "counters", make_ab_copy(env, t->counters, 32, napi_uint32_array, 8));
    set_val(env, o, "secStats", make_ab_copy(env, t->sec_stats, 32, napi_uint32_array, 8));
    set_val(env, o, "dbg", make_ab_copy(env, t->dbg, 32, napi_uint32_array, 8));
    set_num(env, o, "doseSum", (double)t->dose_sum);
    set_num(env, o, "radNRaw", (double)t->rad_n_raw);
    set_num(env, o, "radN", (double)t->rad_n_stored);
    set_num(env, o, "radDropped", (double)t->rad_dropped);
    set_num(env, o, "secNRaw", (double)t->sec_n_raw);
    set_num(env, o, "secN", (double)t->sec_n);
    set_num(env, o, "secDropped", (double)t->sec_dropped);
    set_num(env, o, "tPrimMs", t->t_prim_ms);
    set_num(env, o, "tSecMs", t->t_sec_ms);
    set_num(env, o, "tTotalMs", t->t_total_ms);
    return o;
}

static napi_value js_run_transport(napi_env env, napi_callback_info info)
{
    size_t argc = 4;
    napi_value argv[4];
    CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    if (argc < 2) { napi_throw_type_error(env, NULL, "runTransport(ctx, params, energies?, copyFlags?)"); return NULL; }
    Handle *h = get_handle(env, argv[0]);
    if (!h) return NULL;
    Req *r = (Req *)calloc(1, sizeof(Req));
    r->h = h;
    napi_create_reference(env, argv[0], 1, &r->hold);   /* the context outlives the request */
    napi_value P = argv[1];
    r->p.n = (uint32_t)get_num(env, P, "n", 0);
    r->p.box = (float)get_num(env, P, "box", 15000);
    r->p.ce = (float)get_num(env, P, "ce", 7.4);
    r->p.ms = (uint32_t)get_num(env, P, "ms", 65536);
    r->p.be = (float)get_num(env, P, "be", 0);
    r->p.max_sec = (uint32_t)get_num(env, P, "maxSec", 0);
    r->p.vc = (uint32_t)get_num(env, P, "vc", 128);
    r->p.max_rad = (uint32_t)get_num(env, P, "maxRad", 0);
    r->p.start_half = (float)get_num(env, P, "startHalf", 0);
    r->p.dna_enable = (uint32_t)get_num(env, P, "dnaEnable", 0);
    r->p.dna_grid_n = (uint32_t)get_num(env, P, "dnaGridN", 0);
    r->p.dir_mode = (uint32_t)get_num(env, P, "dirMode", 0);
    r->p.dna_rise = (float)get_num(env, P, "dnaRise", 0);
    r->p.dna_spacing = (float)get_num(env, P, "dnaSpacing", 0);
    r->p.dna_x0 = (float)get_num(env, P, "dnaX0", 0);
    r->p.dna_r_bb = (float)get_num(env, P, "dnaRbb", 0);
    r->p.seed = (uint32_t)get_num(env, P, "seed", 0);
    r->p.max_sec_steps = (uint32_t)get_num(env, P, "maxSecSteps", 2000);
    size_t ne = 0;
    if (argc > 2 && get_typed(env, argv[2], napi_float32_array, (void **)&r->energies, &ne, &r->keep[0])) {
        free_req(env, r);
        napi_throw_type_error(env, NULL, "energies must be a Float32Array or null");
        return NULL;
    }
    if (r->energies && ne < r->p.n) {
        free_req(env, r);
        napi_throw_range_error(env, NULL, "energies has fewer than params.n entries");
        return NULL;
    }
    double cf = DNA_COPY_ALL;
    if (argc > 3) napi_get_value_double(env, argv[3], &cf);
    r->copy = (uint32_t)cf;
    r->exec = exec_transport;
    r->done = done_transport;
    return start(env, r, "dna_run_transport");
}

/* ---------------------------------------------------------------- IRT */
static int exec_irt(Req *r)
{
    return dna_run_irt(r->h->ctx, r->rad, r->rad_count, r->n_therm, r->E, r->tck, r->nck, r->seed, &r->cout);
}

static napi_value done_irt(napi_env env, Req *r)
{
    const dna_chem_out *c = &r->cout;
    napi_value o;
    napi_create_object(env, &o);
    int nc = c->n_ckpt;
    double tally[DNA_MAX_CKPT * 5], G[DNA_MAX_CKPT * 5], rx[DNA_N_RXN];
    for (int k = 0; k < nc; k++)
        for (int s = 0; s < 5; s++) { tally[k * 5 + s] = (double)c->tally[k][s]; G[k * 5 + s] = c->G[k][s]; }
    for (int i = 0; i < DNA_N_RXN; i++) rx[i] = (double)c->rxn_counts[i];
    set_num(env, o, "nCkpt", nc);
    set_val(env, o, "tCkpt", make_ab_copy(env, c->t_ckpt, nc * 8, napi_float64_array, nc));
    set_val(env, o, "tally", make_ab_copy(env, tally, nc * 5 * 8, napi_float64_array, nc * 5));
    set_val(env, o, "G", make_ab_copy(env, G, nc * 5 * 8, napi_float64_array, nc * 5));
    set_val(env, o, "rxnCounts", make_ab_copy(env, rx, DNA_N_RXN * 8, napi_float64_array, DNA_N_RXN));
    set_num(env, o, "nReacted", (double)c->n_reacted);
    set_num(env, o, "chemN", (double)c->chem_n);
    set_num(env, o, "nTracks", (double)c->n_tracks);
    set_num(env, o, "nEvents", (double)c->n_events);
    set_num(env, o, "tMs", c->t_ms);
    set_num(env, o, "tKernelMs", c->t_kernel_ms);
    return o;
}

static napi_value js_run_irt(napi_env env, napi_callback_info info)
{
    size_t argc = 7;
    napi_value argv[7];
    CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    if (argc < 6) { napi_throw_type_error(env, NULL, "runIrt(ctx, rad|null, radN, nTherm, E_eV, tCkpt, seed?)"); return NULL; }
    Handle *h = get_handle(env, argv[0]);
    if (!h) return NULL;
    Req *r = (Req *)calloc(1, sizeof(Req));
    r->h = h;
    napi_create_reference(env, argv[0], 1, &r->hold);   /* the context outlives the request */
    size_t nr = 0, nck = 0;
    double d;
    if (get_typed(env, argv[1], napi_float32_array, (void **)&r->rad, &nr, &r->keep[0]) ||
        get_typed(env, argv[5], napi_float64_array, (void **)&r->tck, &nck, &r->keep[1]) || !r->tck) {
        free_req(env, r);
        napi_throw_type_error(env, NULL, "rad must be a Float32Array or null and tCkpt a Float64Array");
        return NULL;
    }
    napi_get_value_double(env, argv[2], &d); r->rad_count = (uint64_t)d;
    if (r->rad && r->rad_count * 4 > nr) {
        free_req(env, r);
        napi_throw_range_error(env, NULL, "radN exceeds the Float32Array length / 4");
        return NULL;
    }
    napi_get_value_double(env, argv[3], &d); r->n_therm = (uint32_t)d;
    napi_get_value_double(env, argv[4], &d); r->E = (float)d;
    r->nck = (int)nck;
    r->seed = 1;
    if (argc > 6) { napi_get_value_double(env, argv[6], &d); r->seed = (uint64_t)d; }
    r->exec = exec_irt;
    r->done = done_irt;
    return start(env, r, "dna_run_irt");
}

/* ---------------------------------------------------------------- scoring */
/* the DNATarget object of buildDNATarget (src/physics/dna-geometry.ts) into r->tgt; its
 * four arrays are kept alive in r->keep[0..3] */
static int parse_target(napi_env env, napi_value T, Req *r)
{
    r->tgt.rise = get_num(env, T, "rise", 0.34);
    r->tgt.r_bb = get_num(env, T, "r_bb", 1.0);
    r->tgt.spacing_nm = get_num(env, T, "spacing_nm", 150);
    r->tgt.x0 = get_num(env, T, "x0", 0);
    r->tgt.L_nm = get_num(env, T, "L_nm", 3000);
    r->tgt.n_bp_per = (uint32_t)get_num(env, T, "n_bp_per", 0);
    r->tgt.n_fibers = (uint32_t)get_num(env, T, "n_fibers", 0);
    r->tgt.n_bp = (uint32_t)get_num(env, T, "n_bp", 0);
    r->tgt.grid_N = (uint32_t)get_num(env, T, "grid_N", 0);
    const char *names[4] = {"fy", "fz", "rbb0", "rbb1"};
    const float **dst[4] = {&r->tgt.fy, &r->tgt.fz, &r->tgt.rbb0, &r->tgt.rbb1};
    size_t need[4] = {r->tgt.n_fibers, r->tgt.n_fibers, 2 * (size_t)r->tgt.n_bp_per, 2 * (size_t)r->tgt.n_bp_per};
    for (int i = 0; i < 4; i++) {
        napi_value a;
        size_t n = 0;
        if (napi_get_named_property(env, T, names[i], &a) != napi_ok ||
            get_typed(env, a, napi_float32_array, (void **)dst[i], &n, &r->keep[i]) || !*dst[i] || n < need[i])
            return -1;
    }
    return 0;
}

static int exec_score(Req *r)
{
    int rc = dna_score(r->h->ctx, &r->tgt, r->rad, r->rad_count, &r->lcg, r->want_hits, &r->sout);
    if (rc) return rc;
    if (r->want_hits && r->sout.hits) {
        r->hits_n = 2 * (size_t)r->tgt.n_bp;
        r->hits_copy = (uint8_t *)malloc(r->hits_n + 1);
        memcpy(r->hits_copy, r->sout.hits, r->hits_n);
    }
    return DNA_OK;
}

static napi_value done_score(napi_env env, Req *r)
{
    const dna_score_out *s = &r->sout;
    napi_value o;
    napi_create_object(env, &o);
    set_num(env, o, "ssbCount", (double)s->ssb_count);
    set_num(env, o, "candidates", (double)s->candidates);
    set_num(env, o, "inReach", (double)s->in_reach);
    set_num(env, o, "dsb", (double)s->dsb);
    set_num(env, o, "ssb0", (double)s->ssb0);
    set_num(env, o, "ssb1", (double)s->ssb1);
    set_num(env, o, "dsbExpected", s->dsb_expected);
    set_num(env, o, "lcgState", (double)r->lcg);
    set_num(env, o, "tMs", s->t_ms);
    set_val(env, o, "hits", r->hits_copy ? make_ab_copy(env, r->hits_copy, r->hits_n, napi_uint8_array, r->hits_n) : NULL);
    return o;
}

static napi_value js_score(napi_env env, napi_callback_info info)
{
    size_t argc = 6;
    napi_value argv[6];
    CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    if (argc < 5) { napi_throw_type_error(env, NULL, "score(ctx, target, rad|null, radN, lcgState, wantHits?)"); return NULL; }
    Handle *h = get_handle(env, argv[0]);
    if (!h) return NULL;
    Req *r = (Req *)calloc(1, sizeof(Req));
    r->h = h;
    napi_create_reference(env, argv[0], 1, &r->hold);   /* the context outlives the request */
    if (parse_target(env, argv[1], r)) {
        free_req(env, r);
        napi_throw_type_error(env, NULL, "target.fy/fz/rbb0/rbb1 must be Float32Arrays of the target's sizes");
        return NULL;
    }
    double d;
    napi_get_value_double(env, argv[3], &d); r->rad_count = (uint64_t)d;
    size_t nr = 0;
    if (get_typed(env, argv[2], napi_float32_array, (void **)&r->rad, &nr, &r->keep[4]) ||
        (r->rad && r->rad_count * 4 > nr)) {
        free_req(env, r);
        napi_throw_type_error(env, NULL, "rad must be a Float32Array of at least 4*radN floats, or null");
        return NULL;
    }
    napi_get_value_double(env, argv[4], &d); r->lcg = (uint32_t)d;
    r->want_hits = 0;
    if (argc > 5) { bool b = false; napi_get_value_bool(env, argv[5], &b); r->want_hits = b; }
    r->exec = exec_score;
    r->done = done_score;
    return start(env, r, "dna_score");
}

/* ---------------------------------------------------------------- hash */
static int exec_hash(Req *r)
{
    size_t nb = ((size_t)1 << r->bits) + 1;
    r->key = (uint32_t *)malloc(r->npos * 4 + 4);
    r->start = (uint32_t *)malloc(nb * 4);
    r->order = (uint32_t *)malloc(r->npos * 4 + 4);
    return dna_hash_cells(r->h->ctx, r->pos, r->npos, r->cell, r->bits, r->key, r->start, r->order);
}

static napi_value done_hash(napi_env env, Req *r)
{
    napi_value o;
    size_t nb = ((size_t)1 << r->bits) + 1;
    napi_create_object(env, &o);
    set_val(env, o, "key", make_ab_copy(env, r->key, r->npos * 4, napi_uint32_array, r->npos));
    set_val(env, o, "cellStart", make_ab_copy(env, r->start, nb * 4, napi_uint32_array, nb));
    set_val(env, o, "order", make_ab_copy(env, r->order, r->npos * 4, napi_uint32_array, r->npos));
    return o;
}

static napi_value js_hash_cells(napi_env env, napi_callback_info info)
{
    size_t argc = 4;
    napi_value argv[4];
    CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    if (argc < 4) { napi_throw_type_error(env, NULL, "hashCells(ctx, pos4, cellNm, hashBits)"); return NULL; }
    Handle *h = get_handle(env, argv[0]);
    if (!h) return NULL;
    Req *r = (Req *)calloc(1, sizeof(Req));
    r->h = h;
    napi_create_reference(env, argv[0], 1, &r->hold);   /* the context outlives the request */
    size_t n = 0;
    if (get_typed(env, argv[1], napi_float32_array, (void **)&r->pos, &n, &r->keep[0]) || !r->pos) {
        free_req(env, r);
        napi_throw_type_error(env, NULL, "pos4 must be a Float32Array");
        return NULL;
    }
    double d;
    r->npos = n / 4;
    napi_get_value_double(env, argv[2], &d); r->cell = (float)d;
    napi_get_value_double(env, argv[3], &d); r->bits = (uint32_t)d;
    if (r->bits < 1 || r->bits > 24) {
        free_req(env, r);
        napi_throw_range_error(env, NULL, "hashBits must be in 1..24");
        return NULL;
    }
    r->exec = exec_hash;
    r->done = done_hash;
    return start(env, r, "dna_hash_cells");
}

/* ---------------------------------------------------------------- options */
static napi_value js_set_option(napi_env env, napi_callback_info info)
{
    size_t argc = 3;
    napi_value argv[3];
    CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    if (argc < 3) { napi_throw_type_error(env, NULL, "setOption(ctx, option, value)"); return NULL; }
    Handle *h = get_handle(env, argv[0]);
    if (!h) return NULL;
    int32_t opt = 0, val = 0;
    napi_get_value_int32(env, argv[1], &opt);
    napi_get_value_int32(env, argv[2], &val);
    pthread_mutex_lock(&h->mu);
    int rc = dna_set_option(h->ctx, opt, val);
    pthread_mutex_unlock(&h->mu);
    if (rc != DNA_OK) { napi_throw_range_error(env, NULL, dna_last_error()); return NULL; }
    return NULL;
}

static napi_value js_get_option(napi_env env, napi_callback_info info)
{
    size_t argc = 2;
    napi_value argv[2], out;
    CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    if (argc < 2) { napi_throw_type_error(env, NULL, "getOption(ctx, option)"); return NULL; }
    Handle *h = get_handle(env, argv[0]);
    if (!h) return NULL;
    int32_t opt = 0;
    int val = 0;
    napi_get_value_int32(env, argv[1], &opt);
    pthread_mutex_lock(&h->mu);
    int rc = dna_get_option(h->ctx, opt, &val);
    pthread_mutex_unlock(&h->mu);
    if (rc != DNA_OK) { napi_throw_range_error(env, NULL, dna_last_error()); return NULL; }
    CHECK(env, napi_create_int32(env, val, &out));
    return out;
}

static napi_value js_irt_reactions(napi_env env, napi_callback_info info)
{
    (void)info;
    double t[54];
    if (dna_irt_reactions(t) != DNA_OK) { napi_throw_error(env, NULL, dna_last_error()); return NULL; }
    return make_ab_copy(env, t, sizeof(t), napi_float64_array, 54);
}

/* ---------------------------------------------------------------- final IRT state */
static int exec_oh_final(Req *r)
{
    int rc = dna_irt_oh_final(r->h->ctx, NULL, 0, &r->n_out);
    if (rc || !r->want_copy) return rc;
    r->oh_copy = (float *)malloc(r->n_out * 16 + 16);
    return dna_irt_oh_final(r->h->ctx, r->oh_copy, r->n_out, &r->n_out);
}

static napi_value done_oh_final(napi_env env, Req *r)
{
    napi_value o;
    napi_create_object(env, &o);
    set_num(env, o, "n", (double)r->n_out);
    set_val(env, o, "oh", r->oh_copy ? make_ab_copy(env, r->oh_copy, r->n_out * 16, napi_float32_array, r->n_out * 4) : NULL);
    return o;
}

static napi_value js_irt_oh_final(napi_env env, napi_callback_info info)
{
    size_t argc = 2;
    napi_value argv[2];
    CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    if (argc < 1) { napi_throw_type_error(env, NULL, "irtOhFinal(ctx, copy?)"); return NULL; }
    Handle *h = get_handle(env, argv[0]);
    if (!h) return NULL;
    Req *r = (Req *)calloc(1, sizeof(Req));
    r->h = h;
    napi_create_reference(env, argv[0], 1, &r->hold);   /* the context outlives the request */
    r->want_copy = 1;
    if (argc > 1) { bool b = true; if (napi_get_value_bool(env, argv[1], &b) == napi_ok) r->want_copy = b; }
    r->exec = exec_oh_final;
    r->done = done_oh_final;
    return start(env, r, "dna_irt_oh_final");
}

static int exec_particles(Req *r)
{
    int rc = dna_irt_particles(r->h->ctx, NULL, 0, &r->n_out);
    if (rc) return rc;
    r->parts_copy = (double *)malloc(r->n_out * 48 + 48);
    return dna_irt_particles(r->h->ctx, r->parts_copy, r->n_out, &r->n_out);
}

static napi_value done_particles(napi_env env, Req *r)
{
    napi_value o;
    napi_create_object(env, &o);
    set_num(env, o, "n", (double)r->n_out);
    set_val(env, o, "particles", make_ab_copy(env, r->parts_copy, r->n_out * 48, napi_float64_array, r->n_out * 6));
    return o;
}

static napi_value js_irt_particles(napi_env env, napi_callback_info info)
{
    size_t argc = 1;
    napi_value argv[1];
    CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    if (argc < 1) { napi_throw_type_error(env, NULL, "irtParticles(ctx)"); return NULL; }
    Handle *h = get_handle(env, argv[0]);
    if (!h) return NULL;
    Req *r = (Req *)calloc(1, sizeof(Req));
    r->h = h;
    napi_create_reference(env, argv[0], 1, &r->hold);   /* the context outlives the request */
    r->exec = exec_particles;
    r->done = done_particles;
    return start(env, r, "dna_irt_particles");
}

/* ---------------------------------------------------------------- indirect + direct scoring */
static int exec_score_oh(Req *r)
{
    int rc = dna_score_oh(r->h->ctx, &r->tgt, r->rad, r->rad_count, r->ohin, r->oh_count, &r->lcg,
                          r->want_hits, &r->sout, &r->ind);
    if (rc) return rc;
    if (r->want_hits && r->sout.hits) {
        r->hits_n = 2 * (size_t)r->tgt.n_bp;
        r->hits_copy = (uint8_t *)malloc(r->hits_n + 1);
        memcpy(r->hits_copy, r->sout.hits, r->hits_n);
    }
    return DNA_OK;
}

static napi_value done_score_oh(napi_env env, Req *r)
{
    napi_value o = done_score(env, r), ind;
    napi_create_object(env, &ind);
    set_num(env, ind, "ssb0", (double)r->ind.ssb0);
    set_num(env, ind, "ssb1", (double)r->ind.ssb1);
    set_num(env, ind, "candidates", (double)r->ind.candidates);
    set_num(env, ind, "inReach", (double)r->ind.in_reach);
    set_val(env, o, "indirect", ind);
    return o;
}

static napi_value js_score_oh(napi_env env, napi_callback_info info)
{
    size_t argc = 8;
    napi_value argv[8];
    CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    if (argc < 7) {
        napi_throw_type_error(env, NULL, "scoreOh(ctx, target, rad|null, radN, oh|null, nOh, lcgState, wantHits?)");
        return NULL;
    }
    Handle *h = get_handle(env, argv[0]);
    if (!h) return NULL;
    Req *r = (Req *)calloc(1, sizeof(Req));
    r->h = h;
    napi_create_reference(env, argv[0], 1, &r->hold);   /* the context outlives the request */
    if (parse_target(env, argv[1], r)) {
        free_req(env, r);
        napi_throw_type_error(env, NULL, "target.fy/fz/rbb0/rbb1 must be Float32Arrays of the target's sizes");
        return NULL;
    }
    double d;
    size_t nr = 0, no = 0;
    napi_get_value_double(env, argv[3], &d); r->rad_count = (uint64_t)d;
    napi_get_value_double(env, argv[5], &d); r->oh_count = (uint64_t)d;
    if (get_typed(env, argv[2], napi_float32_array, (void **)&r->rad, &nr, &r->keep[4]) ||
        (r->rad && r->rad_count * 4 > nr) ||
        get_typed(env, argv[4], napi_float32_array, (void **)&r->ohin, &no, &r->keep[5]) ||
        (r->ohin && r->oh_count * 4 > no)) {
        free_req(env, r);
        napi_throw_type_error(env, NULL, "rad / oh must be Float32Arrays of at least 4*radN / 4*nOh floats, or null");
        return NULL;
    }
    napi_get_value_double(env, argv[6], &d); r->lcg = (uint32_t)d;
    r->want_hits = 0;
    if (argc > 7) { bool b = false; napi_get_value_bool(env, argv[7], &b); r->want_hits = b; }
    r->exec = exec_score_oh;
    r->done = done_score_oh;
    return start(env, r, "dna_score_oh");
}

/* ---------------------------------------------------------------- module */
static napi_value init(napi_env env, napi_value exports)
{
    napi_property_descriptor d[] = {
        {"version", NULL, js_version, NULL, NULL, NULL, napi_enumerable, NULL},
        {"lastError", NULL, js_last_error, NULL, NULL, NULL, napi_enumerable, NULL},
        {"ctxCreate", NULL, js_ctx_create, NULL, NULL, NULL, napi_enumerable, NULL},
        {"ctxDestroy", NULL, js_ctx_destroy, NULL, NULL, NULL, napi_enumerable, NULL},
        {"runTransport", NULL, js_run_transport, NULL, NULL, NULL, napi_enumerable, NULL},
        {"runIrt", NULL, js_run_irt, NULL, NULL, NULL, napi_enumerable, NULL},
        {"score", NULL, js_score, NULL, NULL, NULL, napi_enumerable, NULL},
        {"hashCells", NULL, js_hash_cells, NULL, NULL, NULL, napi_enumerable, NULL},
        {"setOption", NULL, js_set_option, NULL, NULL, NULL, napi_enumerable, NULL},
        {"getOption", NULL, js_get_option, NULL, NULL, NULL, napi_enumerable, NULL},
        {"irtOhFinal", NULL, js_irt_oh_final, NULL, NULL, NULL, napi_enumerable, NULL},
        {"irtParticles", NULL, js_irt_particles, NULL, NULL, NULL, napi_enumerable, NULL},
        {"scoreOh", NULL, js_score_oh, NULL, NULL, NULL, napi_enumerable, NULL},
        {"irtReactions", NULL, js_irt_reactions, NULL, NULL, NULL, napi_enumerable, NULL},
    };
    CHECK(env, napi_define_properties(env, exports, sizeof(d) / sizeof(d[0]), d));
    return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)
