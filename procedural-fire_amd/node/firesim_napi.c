/*
 * N-API addon binding libfiresim.so (include/firesim.h) for Node.js.
 *
 * A thin layer: every function converts JS values to the C ABI's plain pointers
 * and sizes and throws a JS Error carrying fs_last_error() when a call fails —
 * the reference's JS FireSim throws on unknown uniforms (src/gpu.js:65).  The
 * handle is a JS external whose finaliser calls fs_destroy.  stepAsync runs
 * fs_step + fs_sync on a libuv worker and resolves a Promise, the counterpart of
 * FIRE.step awaiting queue.onSubmittedWorkDone() (src/main.js:244).
 */
#include <node_api.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/firesim.h"

#define CHECK_NAPI(env, call)                                              \
    do {                                                                   \
        if ((call) != napi_ok) {                                           \
            napi_throw_error((env), NULL, "N-API call failed: " #call);    \
            return NULL;                                                   \
        }                                                                  \
    } while (0)

static napi_value throw_fs(napi_env env, int rc) {
    char msg[512];
    snprintf(msg, sizeof msg, "firesim error %d: %s", rc, fs_last_error());
    napi_throw_error(env, NULL, msg);
    return NULL;
}

static void finalize_sim(napi_env env, void* data, void* hint) {
    (void)env; (void)hint;
    fs_destroy((fs_sim*)data);
}

static fs_sim* get_sim(napi_env env, napi_value v) {
    void* p = NULL;
    if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
        napi_throw_type_error(env, NULL, "expected a FireSim handle");
        return NULL;
    }
    return (fs_sim*)p;
}

static int get_prop_num(napi_env env, napi_value obj, const char* key, double* out) {
    bool has = false;
    if (napi_has_named_property(env, obj, key, &has) != napi_ok || !has) return 0;
    napi_value v;
    napi_get_named_property(env, obj, key, &v);
    napi_valuetype t;
    napi_typeof(env, v, &t);
    if (t != napi_number) return 0;
    napi_get_value_double(env, v, out);
    return 1;
}

static int get_prop_array3(napi_env env, napi_value obj, const char* key, double out[3]) {
    bool has = false;
    if (napi_has_named_property(env, obj, key, &has) != napi_ok || !has) return 0;
    napi_value v;
    napi_get_named_property(env, obj, key, &v);
    bool isarr = false;
    napi_is_array(env, v, &isarr);
    if (!isarr) return 0;
    for (uint32_t i = 0; i < 3; i++) {
        napi_value e;
        napi_get_element(env, v, i, &e);
        napi_get_value_double(env, e, &out[i]);
    }
    return 1;
}

/* typed array -> data pointer + byte length */
static int typed_data(napi_env env, napi_value v, void** data, size_t* bytes, napi_typedarray_type* type) {
    bool is = false;
    napi_is_typedarray(env, v, &is);
    if (!is) return 0;
    size_t len, off;
    napi_value ab;
    napi_get_typedarray_info(env, v, type, &len, data, &ab, &off);
    size_t el = 1;
    switch (*type) {
        case napi_float32_array: case napi_int32_array: case napi_uint32_array: el = 4; break;
        case napi_float64_array: el = 8; break;
        case napi_uint16_array: case napi_int16_array: el = 2; break;
        default: el = 1;
    }
    *bytes = len * el;
    return 1;
}

/* create(opts): opts = {res, size, logs(Float32Array), vcycles, maxLevels, precision, solve,
 *                       divTol, maxVcycles, noiseN, noiseSeed, device, useGraph, adaptive, omega} */
static napi_value js_create(napi_env env, napi_callback_info info) {
    size_t argc = 1;
    napi_value argv[1];
    CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    fs_config cfg;
    fs_config_default(&cfg);
    const float* logs = NULL;
    int n_logs = 0;
    if (argc >= 1) {
        napi_valuetype t;
        napi_typeof(env, argv[0], &t);
        if (t == napi_object) {
            double a3[3], d;
            if (get_prop_array3(env, argv[0], "res", a3)) for (int i = 0; i < 3; i++) cfg.res[i] = (int)a3[i];
            if (get_prop_array3(env, argv[0], "size", a3)) for (int i = 0; i < 3; i++) cfg.size[i] = (float)a3[i];
            if (get_prop_num(env, argv[0], "vcycles", &d)) cfg.vcycles = (int)d;
            if (get_prop_num(env, argv[0], "maxLevels", &d)) cfg.max_levels = (int)d;
            if (get_prop_num(env, argv[0], "precision", &d)) cfg.precision = (int)d;
            if (get_prop_num(env, argv[0], "solve", &d)) cfg.solve = (int)d;
            if (get_prop_num(env, argv[0], "divTol", &d)) cfg.div_tol = d;
            if (get_prop_num(env, argv[0], "maxVcycles", &d)) cfg.max_vcycles = (int)d;
            if (get_prop_num(env, argv[0], "noiseN", &d)) cfg.noise_n = (int)d;
            if (get_prop_num(env, argv[0], "noiseSeed", &d)) cfg.noise_seed = (int)d;
            if (get_prop_num(env, argv[0], "device", &d)) cfg.device = (int)d;
            if (get_prop_num(env, argv[0], "useGraph", &d)) cfg.use_graph = (int)d;
            if (get_prop_num(env, argv[0], "adaptive", &d)) cfg.adaptive = (int)d;
            if (get_prop_num(env, argv[0], "omega", &d)) cfg.omega = (float)d;
            if (get_prop_num(env, argv[0], "preSweeps", &d)) cfg.pre_sweeps = (int)d;
            if (get_prop_num(env, argv[0], "postSweeps", &d)) cfg.post_sweeps = (int)d;
            if (get_prop_num(env, argv[0], "directLevel", &d)) cfg.direct_level = (int)d;
            if (get_prop_num(env, argv[0], "firstCycles", &d)) cfg.first_cycles = (int)d;
            bool has = false;
            napi_has_named_property(env, argv[0], "logs", &has);
            if (has) {
                napi_value lv;
                napi_get_named_property(env, argv[0], "logs", &lv);
                void* data;
                size_t bytes;
                napi_typedarray_type tt;
                if (!typed_data(env, lv, &data, &bytes, &tt) || tt != napi_float32_array) {
                    napi_throw_type_error(env, NULL, "logs must be a Float32Array (packLogs layout)");
                    return NULL;
                }
                logs = (const float*)data;
                n_logs = (int)(bytes / 32);
            }
        }
    }
    fs_sim* h = NULL;
    int rc = fs_create(&cfg, logs, n_logs, &h);
    if (rc) return throw_fs(env, rc);
    napi_value ext;
    CHECK_NAPI(env, napi_create_external(env, h, finalize_sim, NULL, &ext));
    return ext;
}

static napi_value js_destroy(napi_env env, napi_callback_info info) {
    /* the handle's finaliser owns destruction; nothing to do eagerly except sync */
    size_t argc = 1;
    napi_value argv[1];
    CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    fs_sim* h = get_sim(env, argv[0]);
    if (!h) return NULL;
    int rc = fs_sync(h);
    if (rc) return throw_fs(env, rc);
    return NULL;
}

static napi_value js_param_names(napi_env env, napi_callback_info info) {
    (void)info;
    napi_value arr;
    CHECK_NAPI(env, napi_create_array(env, &arr));
    for (int i = 0; i < fs_param_count(); i++) {
        napi_value s;
        napi_create_string_utf8(env, fs_param_name(i), NAPI_AUTO_LENGTH, &s);
        napi_set_element(env, arr, (uint32_t)i, s);
    }
    return arr;
}

static napi_value float32_array(napi_env env, const float* src, size_t n) {
    void* data;
    napi_value ab, ta;
    if (napi_create_arraybuffer(env, n * sizeof(float), &data, &ab) != napi_ok) return NULL;
    memcpy(data, src, n * sizeof(float));
    napi_create_typedarray(env, napi_float32_array, n, ab, 0, &ta);
    return ta;
}

static napi_value js_param_defaults(napi_env env, napi_callback_info info) {
    (void)info;
    float p[FS_NUM_PARAMS];
    fs_param_defaults(p);
    return float32_array(env, p, FS_NUM_PARAMS);
}

/* setParams(h, names: string[], values: number[]) */
static napi_value js_set_params(napi_env env, napi_callback_info info) {
    size_t argc = 3;
    napi_value argv[3];
    CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    fs_sim* h = get_sim(env, argv[0]);
    if (!h) return NULL;
    uint32_t n = 0;
    napi_get_array_length(env, argv[1], &n);
    char** names = (char**)calloc(n ? n : 1, sizeof(char*));
    float* vals = (float*)calloc(n ? n : 1, sizeof(float));
    for (uint32_t i = 0; i < n; i++) {
        napi_value s, v;
        napi_get_element(env, argv[1], i, &s);
        napi_get_element(env, argv[2], i, &v);
        size_t len = 0;
        napi_get_value_string_utf8(env, s, NULL, 0, &len);
        names[i] = (char*)malloc(len + 1);
        napi_get_value_string_utf8(env, s, names[i], len + 1, &len);
        double d = 0;
        napi_get_value_double(env, v, &d);
        vals[i] = (float)d;
    }
    int rc = fs_set_params(h, (const char* const*)names, vals, (int)n);
    for (uint32_t i = 0; i < n; i++) free(names[i]);
    free(names);
    free(vals);
    if (rc) return throw_fs(env, rc);
    return NULL;
}

static napi_value js_get_params(napi_env env, napi_callback_info info) {
    size_t argc = 1;
    napi_value argv[1];
    CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    fs_sim* h = get_sim(env, argv[0]);
    if (!h) return NULL;
    float p[FS_NUM_PARAMS];
    int rc = fs_get_params(h, p);
    if (rc) return throw_fs(env, rc);
    return float32_array(env, p, FS_NUM_PARAMS);
}

static napi_value js_set_logs(napi_env env, napi_callback_info info) {
    size_t argc = 2;
    napi_value argv[2];
    CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    fs_sim* h = get_sim(env, argv[0]);
    if (!h) return NULL;
    void* data;
    size_t bytes;
    napi_typedarray_type t;
    if (!typed_data(env, argv[1], &data, &bytes, &t) || t != napi_float32_array) {
        napi_throw_type_error(env, NULL, "logs must be a Float32Array");
        return NULL;
    }
    int rc = fs_set_logs(h, (const float*)data, (int)(bytes / 32));
    if (rc) return throw_fs(env, rc);
    return NULL;
}

/* setSpots(h, Float32Array of n*6: x, y, z, radius, flux_kg_s, temp) -> Float32Array of n normalisations */
static napi_value js_set_spots(napi_env env, napi_callback_info info) {
    size_t argc = 2;
    napi_value argv[2];
    CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    fs_sim* h = get_sim(env, argv[0]);
    if (!h) return NULL;
    void* data;
    size_t bytes;
    napi_typedarray_type t;
    if (!typed_data(env, argv[1], &data, &bytes, &t) || t != napi_float32_array || (bytes / 4) % 6) {
        napi_throw_type_error(env, NULL, "spots must be a Float32Array of 6 floats per spot");
        return NULL;
    }
    const int n = (int)(bytes / 24);
    if (n > FS_MAX_SPOTS) {
        napi_throw_range_error(env, NULL, "too many spots");
        return NULL;
    }
    fs_spot sp[FS_MAX_SPOTS];
    const float* f = (const float*)data;
    for (int i = 0; i < n; i++) {
        memset(&sp[i], 0, sizeof(fs_spot));
        sp[i].pos[0] = f[6 * i]; sp[i].pos[1] = f[6 * i + 1]; sp[i].pos[2] = f[6 * i + 2];
        sp[i].radius = f[6 * i + 3]; sp[i].flux = f[6 * i + 4]; sp[i].temp = f[6 * i + 5];
    }
    int rc = fs_set_spots(h, sp, n);
    if (rc) return throw_fs(env, rc);
    int got = 0;
    rc = fs_get_spots(h, sp, FS_MAX_SPOTS, &got);
    if (rc) return throw_fs(env, rc);
    float norm[FS_MAX_SPOTS];
    for (int i = 0; i < got; i++) norm[i] = sp[i].norm;
    return float32_array(env, norm, (size_t)got);
}

/* setFaceFlux(h, Float32Array nx*nz | null) */
static napi_value js_set_face_flux(napi_env env, napi_callback_info info) {
    size_t argc = 2;
    napi_value argv[2];
    CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    fs_sim* h = get_sim(env, argv[0]);
    if (!h) return NULL;
    napi_valuetype vt;
    CHECK_NAPI(env, napi_typeof(env, argv[1], &vt));
    if (vt == napi_null || vt == napi_undefined) {
        int rc = fs_set_face_flux(h, NULL);
        if (rc) return throw_fs(env, rc);
        return NULL;
    }
    void* data;
    size_t bytes;
    napi_typedarray_type t;
    int res[3];
    double dx, origin[3];
    int rc = fs_grid_info(h, res, &dx, origin);
    if (rc) return throw_fs(env, rc);
    if (!typed_data(env, argv[1], &data, &bytes, &t) || t != napi_float32_array ||
        bytes != (size_t)res[0] * res[2] * sizeof(float)) {
        napi_throw_type_error(env, NULL, "face flux must be a Float32Array of nx*nz floats");
        return NULL;
    }
    rc = fs_set_face_flux(h, (const float*)data);
    if (rc) return throw_fs(env, rc);
    return NULL;
}

static napi_value js_reset(napi_env env, napi_callback_info info) {
    size_t argc = 1;
    napi_value argv[1];
    CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    fs_sim* h = get_sim(env, argv[0]);
    if (!h) return NULL;
    int rc = fs_reset(h);
    if (rc) return throw_fs(env, rc);
    return NULL;
}

/* step(h, dt, n): enqueue only (like encoding into a command buffer) */
static napi_value js_step(napi_env env, napi_callback_info info) {
    size_t argc = 3;
    napi_value argv[3];
    CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    fs_sim* h = get_sim(env, argv[0]);
    if (!h) return NULL;
    double dt = 0, n = 1;
    napi_get_value_double(env, argv[1], &dt);
    if (argc >= 3) napi_get_value_double(env, argv[2], &n);
    int rc = fs_step(h, dt, (int)n);
    if (rc) return throw_fs(env, rc);
    return NULL;
}

static napi_value js_sync(napi_env env, napi_callback_info info) {
    size_t argc = 1;
    napi_value argv[1];
    CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    fs_sim* h = get_sim(env, argv[0]);
    if (!h) return NULL;
    int rc = fs_sync(h);
    if (rc) return throw_fs(env, rc);
    return NULL;
}

/* stepAsync(h, dt, n) -> Promise: fs_step + fs_sync on a worker thread */
typedef struct {
    napi_async_work work;
    napi_deferred deferred;
    napi_ref handle_ref;
    fs_sim* h;
    double dt;
    int n;
    int rc;
    char err[512];
} StepWork;

static void step_execute(napi_env env, void* data) {
    (void)env;
    StepWork* w = (StepWork*)data;
    w->rc = fs_step(w->h, w->dt, w->n);
    if (!w->rc) w->rc = fs_sync(w->h);
    if (w->rc) snprintf(w->err, sizeof w->err, "firesim error %d: %s", w->rc, fs_last_error());
}

static void step_complete(napi_env env, napi_status status, void* data) {
    StepWork* w = (StepWork*)data;
    napi_value undef;
    napi_get_undefined(env, &undef);
    if (status != napi_ok || w->rc) {
        napi_value msg, err;
        napi_create_string_utf8(env, w->rc ? w->err : "async step cancelled", NAPI_AUTO_LENGTH, &msg);
        napi_create_error(env, NULL, msg, &err);
        napi_reject_deferred(env, w->deferred, err);
    } else {
        napi_resolve_deferred(env, w->deferred, undef);
    }
    napi_delete_reference(env, w->handle_ref);
    napi_delete_async_work(env, w->work);
    free(w);
}

static napi_value js_step_async(napi_env env, napi_callback_info info) {
    size_t argc = 3;
    napi_value argv[3];
    CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    fs_sim* h = get_sim(env, argv[0]);
    if (!h) return NULL;
    StepWork* w = (StepWork*)calloc(1, sizeof(StepWork));
    w->h = h;
    w->n = 1;
    napi_get_value_double(env, argv[1], &w->dt);
    if (argc >= 3) {
        double n = 1;
        napi_get_value_double(env, argv[2], &n);
        w->n = (int)n;
    }
    napi_value promise, name;
    CHECK_NAPI(env, napi_create_promise(env, &w->deferred, &promise));
    /* keep the handle alive while the worker uses it */
    CHECK_NAPI(env, napi_create_reference(env, argv[0], 1, &w->handle_ref));
    napi_create_string_utf8(env, "firesim.stepAsync", NAPI_AUTO_LENGTH, &name);
    CHECK_NAPI(env, napi_create_async_work(env, NULL, name, step_execute, step_complete, w, &w->work));
    CHECK_NAPI(env, napi_queue_async_work(env, w->work));
    return promise;
}

/* readField(h, field, native: bool) -> Float32Array (or Uint16Array / Float64Array raw storage) */
static napi_value js_read_field(napi_env env, napi_callback_info info) {
    size_t argc = 3;
    napi_value argv[3];
    CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    fs_sim* h = get_sim(env, argv[0]);
    if (!h) return NULL;
    int32_t field = 0;
    napi_get_value_int32(env, argv[1], &field);
    bool native = false;
    if (argc >= 3) napi_get_value_bool(env, argv[2], &native);
    int dims[3], ch = 0, nb = 0;
    int rc = fs_field_info(h, field, dims, &ch, &nb);
    if (rc) return throw_fs(env, rc);
    size_t n = (size_t)dims[0] * dims[1] * dims[2] * ch;
    size_t el = native ? (size_t)nb : 4;
    void* data;
    napi_value ab, ta;
    CHECK_NAPI(env, napi_create_arraybuffer(env, n * el, &data, &ab));
    rc = fs_read_field(h, field, native ? FS_DT_NATIVE : FS_DT_F32, data, n * el);
    if (rc) return throw_fs(env, rc);
    napi_typedarray_type t = !native ? napi_float32_array : nb == 2 ? napi_uint16_array : nb == 8 ? napi_float64_array
                                                                                                 : napi_float32_array;
    CHECK_NAPI(env, napi_create_typedarray(env, t, n, ab, 0, &ta));
    return ta;
}

/* writeField(h, field, typedArray): Float32Array / Float64Array converted, Uint16Array raw */
static napi_value js_write_field(napi_env env, napi_callback_info info) {
    size_t argc = 3;
    napi_value argv[3];
    CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    fs_sim* h = get_sim(env, argv[0]);
    if (!h) return NULL;
    int32_t field = 0;
    napi_get_value_int32(env, argv[1], &field);
    void* data;
    size_t bytes;
    napi_typedarray_type t;
    if (!typed_data(env, argv[2], &data, &bytes, &t)) {
        napi_throw_type_error(env, NULL, "expected a typed array");
        return NULL;
    }
    int dt = t == napi_float32_array ? FS_DT_F32 : t == napi_float64_array ? FS_DT_F64 : FS_DT_NATIVE;
    int rc = fs_write_field(h, field, dt, data, bytes);
    if (rc) return throw_fs(env, rc);
    return NULL;
}

static napi_value js_info(napi_env env, napi_callback_info info) {
    size_t argc = 1;
    napi_value argv[1];
    CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    fs_sim* h = get_sim(env, argv[0]);
    if (!h) return NULL;
    double t = 0, dx = 0, org[3];
    int64_t frame = 0;
    int res[3], levels = 0, coarse = 0, ldims[16][3];
    if (fs_get_clock(h, &t, &frame) || fs_grid_info(h, res, &dx, org) || fs_mg_info(h, &levels, ldims, &coarse))
        return throw_fs(env, FS_EINVAL);
    napi_value o, v, arr;
    napi_create_object(env, &o);
    napi_create_double(env, t, &v); napi_set_named_property(env, o, "time", v);
    napi_create_int64(env, frame, &v); napi_set_named_property(env, o, "frame", v);
    napi_create_double(env, dx, &v); napi_set_named_property(env, o, "dx", v);
    napi_create_array(env, &arr);
    for (int i = 0; i < 3; i++) { napi_create_int32(env, res[i], &v); napi_set_element(env, arr, i, v); }
    napi_set_named_property(env, o, "res", arr);
    napi_create_array(env, &arr);
    for (int i = 0; i < 3; i++) { napi_create_double(env, org[i], &v); napi_set_element(env, arr, i, v); }
    napi_set_named_property(env, o, "origin", arr);
    napi_create_array(env, &arr);
    for (int l = 0; l < levels; l++) {
        napi_value d;
        napi_create_array(env, &d);
        for (int k = 0; k < 3; k++) { napi_create_int32(env, ldims[l][k], &v); napi_set_element(env, d, k, v); }
        napi_set_element(env, arr, l, d);
    }
    napi_set_named_property(env, o, "levels", arr);
    napi_create_int32(env, coarse, &v); napi_set_named_property(env, o, "coarseIters", v);
    return o;
}

static napi_value float64_array(napi_env env, const double* src, size_t n) {
    napi_value ab, ta;
    void* data;
    napi_create_arraybuffer(env, n * sizeof(double), &data, &ab);
    if (n) memcpy(data, src, n * sizeof(double));
    napi_create_typedarray(env, napi_float64_array, n, ab, 0, &ta);
    return ta;
}

/* stats(h, field) -> Float64Array(20): per channel min, max, mean, bad, argmax index */
static napi_value js_stats(napi_env env, napi_callback_info info) {
    size_t argc = 2;
    napi_value argv[2];
    CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    fs_sim* h = get_sim(env, argv[0]);
    if (!h) return NULL;
    int32_t field = 0;
    napi_get_value_int32(env, argv[1], &field);
    double out[20];
    int rc = fs_stats(h, field, out);
    if (rc) return throw_fs(env, rc);
    return float64_array(env, out, 20);
}

/* profileRows(h) -> Float64Array: [Q_W, then 6 doubles per row] */
static napi_value js_profile_rows(napi_env env, napi_callback_info info) {
    size_t argc = 1;
    napi_value argv[1];
    CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    fs_sim* h = get_sim(env, argv[0]);
    if (!h) return NULL;
    int res[3];
    double dx, org[3];
    if (fs_grid_info(h, res, &dx, org)) return throw_fs(env, FS_EINVAL);
    const int max_rows = (res[1] + 3) / 4;
    double* buf = (double*)malloc(sizeof(double) * (1 + 6 * (size_t)max_rows));
    if (!buf) { napi_throw_error(env, NULL, "out of memory"); return NULL; }
    int n = 0;
    int rc = fs_profile(h, buf, buf + 1, max_rows, &n);
    if (rc) { free(buf); return throw_fs(env, rc); }
    napi_value ta = float64_array(env, buf, 1 + 6 * (size_t)n);
    free(buf);
    return ta;
}

/* diagEnable(h, capacity, tFlame) */
static napi_value js_diag_enable(napi_env env, napi_callback_info info) {
    size_t argc = 3;
    napi_value argv[3];
    CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    fs_sim* h = get_sim(env, argv[0]);
    if (!h) return NULL;
    int32_t cap = 0;
    double tf = 1000.0;
    napi_get_value_int32(env, argv[1], &cap);
    if (argc > 2) napi_get_value_double(env, argv[2], &tf);
    int rc = fs_diag_enable(h, cap, tf);
    if (rc) return throw_fs(env, rc);
    return NULL;
}

/* diagRead(h, maxRecords) -> Float64Array of n * (4 + ny) doubles, oldest first */
static napi_value js_diag_read(napi_env env, napi_callback_info info) {
    size_t argc = 2;
    napi_value argv[2];
    CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    fs_sim* h = get_sim(env, argv[0]);
    if (!h) return NULL;
    int32_t maxr = 0;
    napi_get_value_int32(env, argv[1], &maxr);
    int res[3];
    double dx, org[3];
    if (fs_grid_info(h, res, &dx, org)) return throw_fs(env, FS_EINVAL);
    const size_t rec = 4 + (size_t)res[1];
    double* buf = (double*)malloc(sizeof(double) * rec * (size_t)(maxr > 0 ? maxr : 1));
    if (!buf) { napi_throw_error(env, NULL, "out of memory"); return NULL; }
    int n = 0;
    int rc = fs_diag_read(h, buf, maxr, &n);
    if (rc) { free(buf); return throw_fs(env, rc); }
    napi_value ta = float64_array(env, buf, rec * (size_t)n);
    free(buf);
    return ta;
}

static napi_value js_residual(napi_env env, napi_callback_info info) {
    size_t argc = 1;
    napi_value argv[1];
    CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    fs_sim* h = get_sim(env, argv[0]);
    if (!h) return NULL;
    double rel = 0;
    int used = 0;
    int rc = fs_residual(h, &rel, &used);
    if (rc) return throw_fs(env, rc);
    napi_value o, v;
    napi_create_object(env, &o);
    napi_create_double(env, rel, &v); napi_set_named_property(env, o, "rel", v);
    napi_create_int32(env, used, &v); napi_set_named_property(env, o, "vcycles", v);
    return o;
}

static napi_value js_profile(napi_env env, napi_callback_info info) {
    size_t argc = 2;
    napi_value argv[2];
    CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    fs_sim* h = get_sim(env, argv[0]);
    if (!h) return NULL;
    bool on = true;
    if (argc >= 2) napi_get_value_bool(env, argv[1], &on);
    int rc = fs_profile_enable(h, on);
    if (rc) return throw_fs(env, rc);
    return NULL;
}

static napi_value js_timings(napi_env env, napi_callback_info info) {
    size_t argc = 1;
    napi_value argv[1];
    CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    fs_sim* h = get_sim(env, argv[0]);
    if (!h) return NULL;
    float ms[FS_NUM_TIMINGS];
    int rc = fs_timings(h, ms);
    if (rc) return throw_fs(env, rc);
    return float32_array(env, ms, FS_NUM_TIMINGS);
}

static napi_value js_build_campfire(napi_env env, napi_callback_info info) {
    size_t argc = 2;
    napi_value argv[2];
    CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    int32_t seed = 7, maxl = 16;
    if (argc >= 1) napi_get_value_int32(env, argv[0], &seed);
    if (argc >= 2) napi_get_value_int32(env, argv[1], &maxl);
    if (maxl < 0 || maxl > 1024) { napi_throw_range_error(env, NULL, "max logs out of range"); return NULL; }
    float* buf = (float*)calloc((size_t)(maxl ? maxl : 1) * 8, sizeof(float));
    int n = fs_build_campfire(seed, buf, maxl);
    napi_value arr = float32_array(env, buf, (size_t)maxl * 8);
    free(buf);
    if (n < 0) return throw_fs(env, n);
    return arr;
}

/* ---- sparks (src/sparks.js) ---------------------------------------------------- */
static void finalize_sparks(napi_env env, void* data, void* hint) {
    (void)env; (void)hint;
    fs_sparks_destroy((fs_sparks*)data);
}
static fs_sparks* get_sparks(napi_env env, napi_value v) {
    void* p = NULL;
    if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
        napi_throw_type_error(env, NULL, "expected a Sparks handle");
        return NULL;
    }
    return (fs_sparks*)p;
}
/* sparksCreate(h, count) */
static napi_value js_sparks_create(napi_env env, napi_callback_info info) {
    size_t argc = 2;
    napi_value argv[2];
    CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    fs_sim* h = get_sim(env, argv[0]);
    if (!h) return NULL;
    double count = 8192;
    if (argc >= 2) napi_get_value_double(env, argv[1], &count);
    fs_sparks* sp = NULL;
    int rc = fs_sparks_create(h, (int)count, &sp);
    if (rc) return throw_fs(env, rc);
    napi_value ext;
    CHECK_NAPI(env, napi_create_external(env, sp, finalize_sparks, NULL, &ext));
    return ext;
}
static napi_value js_sparks_reset(napi_env env, napi_callback_info info) {
    size_t argc = 1;
    napi_value argv[1];
    CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    fs_sparks* sp = get_sparks(env, argv[0]);
    if (!sp) return NULL;
    int rc = fs_sparks_reset(sp);
    if (rc) return throw_fs(env, rc);
    return NULL;
}
/* sparksSet(s, name, value) */
static napi_value js_sparks_set(napi_env env, napi_callback_info info) {
    size_t argc = 3;
    napi_value argv[3];
    CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    fs_sparks* sp = get_sparks(env, argv[0]);
    if (!sp) return NULL;
    char name[64];
    size_t len = 0;
    if (napi_get_value_string_utf8(env, argv[1], name, sizeof name, &len) != napi_ok) {
        napi_throw_type_error(env, NULL, "expected a parameter name");
        return NULL;
    }
    double v = 0;
    napi_get_value_double(env, argv[2], &v);
    int rc = fs_sparks_set(sp, name, (float)v);
    if (rc) return throw_fs(env, rc);
    return NULL;
}
/* sparksBurst(s, {n, pos, speed, radius, dur, fan, size, aim}) (src/sparks.js:44-46) */
static napi_value js_sparks_burst(napi_env env, napi_callback_info info) {
    size_t argc = 2;
    napi_value argv[2];
    CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    fs_sparks* sp = get_sparks(env, argv[0]);
    if (!sp) return NULL;
    fs_spark_burst b;
    fs_spark_burst_default(&b);
    if (argc >= 2) {
        double v, a3[3];
        if (get_prop_num(env, argv[1], "n", &v)) b.n = (int)v;
        if (get_prop_array3(env, argv[1], "pos", a3)) for (int k = 0; k < 3; k++) b.pos[k] = (float)a3[k];
        if (get_prop_num(env, argv[1], "speed", &v)) b.speed = (float)v;
        if (get_prop_num(env, argv[1], "radius", &v)) b.radius = (float)v;
        if (get_prop_num(env, argv[1], "dur", &v)) b.dur = v;
        if (get_prop_num(env, argv[1], "fan", &v)) b.fan = (float)v;
        if (get_prop_num(env, argv[1], "size", &v)) b.size = (float)v;
        if (get_prop_array3(env, argv[1], "aim", a3)) { b.has_aim = 1; for (int k = 0; k < 3; k++) b.aim[k] = (float)a3[k]; }
    }
    int rc = fs_sparks_burst(sp, &b);
    if (rc) return throw_fs(env, rc);
    return NULL;
}
/* sparksStep(s, dt) */
static napi_value js_sparks_step(napi_env env, napi_callback_info info) {
    size_t argc = 2;
    napi_value argv[2];
    CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    fs_sparks* sp = get_sparks(env, argv[0]);
    if (!sp) return NULL;
    double dt = 0;
    napi_get_value_double(env, argv[1], &dt);
    int rc = fs_sparks_step(sp, dt);
    if (rc) return throw_fs(env, rc);
    return NULL;
}
/* sparksRead(s, count) -> Float32Array(count * 16) */
static napi_value js_sparks_read(napi_env env, napi_callback_info info) {
    size_t argc = 2;
    napi_value argv[2];
    CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    fs_sparks* sp = get_sparks(env, argv[0]);
    if (!sp) return NULL;
    double count = 0;
    napi_get_value_double(env, argv[1], &count);
    const size_t n = (size_t)count * FS_SPARK_FLOATS;
    void* data = NULL;
    napi_value ab, arr;
    CHECK_NAPI(env, napi_create_arraybuffer(env, n * 4, &data, &ab));
    int rc = fs_sparks_read(sp, (float*)data, n * 4);
    if (rc) return throw_fs(env, rc);
    CHECK_NAPI(env, napi_create_typedarray(env, napi_float32_array, n, ab, 0, &arr));
    return arr;
}

/* ---- render-volume consumers (src/renderer.js:104-326) -------------------------- */
static void finalize_render(napi_env env, void* data, void* hint) {
    (void)env; (void)hint;
    fs_render_destroy((fs_render*)data);
}
static fs_render* get_render(napi_env env, napi_value v) {
    void* p = NULL;
    if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
        napi_throw_type_error(env, NULL, "expected a Render handle");
        return NULL;
    }
    return (fs_render*)p;
}
/* renderCreate(h, width, height, volScale) */
static napi_value js_render_create(napi_env env, napi_callback_info info) {
    size_t argc = 4;
    napi_value argv[4];
    CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    fs_sim* h = get_sim(env, argv[0]);
    if (!h) return NULL;
    double w = 960, ht = 1080, vs = 0.5;
    if (argc >= 2) napi_get_value_double(env, argv[1], &w);
    if (argc >= 3) napi_get_value_double(env, argv[2], &ht);
    if (argc >= 4) napi_get_value_double(env, argv[3], &vs);
    fs_render* r = NULL;
    int rc = fs_render_create(h, (int)w, (int)ht, vs, &r);
    if (rc) return throw_fs(env, rc);
    napi_value ext;
    CHECK_NAPI(env, napi_create_external(env, r, finalize_render, NULL, &ext));
    return ext;
}
/* renderSet(r, name, value) */
static napi_value js_render_set(napi_env env, napi_callback_info info) {
    size_t argc = 3;
    napi_value argv[3];
    CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    fs_render* r = get_render(env, argv[0]);
    if (!r) return NULL;
    char name[64];
    size_t len = 0;
    if (napi_get_value_string_utf8(env, argv[1], name, sizeof name, &len) != napi_ok) {
        napi_throw_type_error(env, NULL, "expected a parameter name");
        return NULL;
    }
    double v = 0;
    napi_get_value_double(env, argv[2], &v);
    int rc = fs_render_set(r, name, (float)v);
    if (rc) return throw_fs(env, rc);
    return NULL;
}
/* renderSetCamera(r, {pos, fwd, right, up, fovY}) (src/camera.js lookAt) */
static napi_value js_render_set_camera(napi_env env, napi_callback_info info) {
    size_t argc = 2;
    napi_value argv[2];
    CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    fs_render* r = get_render(env, argv[0]);
    if (!r) return NULL;
    double pos[3], fwd[3], rgt[3], up[3], fov = 0;
    if (argc < 2 || !get_prop_array3(env, argv[1], "pos", pos) || !get_prop_array3(env, argv[1], "fwd", fwd) ||
        !get_prop_array3(env, argv[1], "right", rgt) || !get_prop_array3(env, argv[1], "up", up) ||
        !get_prop_num(env, argv[1], "fovY", &fov)) {
        napi_throw_type_error(env, NULL, "expected a camera {pos, fwd, right, up, fovY}");
        return NULL;
    }
    int rc = fs_render_set_camera(r, pos, fwd, rgt, up, fov);
    if (rc) return throw_fs(env, rc);
    return NULL;
}
/* renderSetWhiteBalance(r, twb) */
static napi_value js_render_set_wb(napi_env env, napi_callback_info info) {
    size_t argc = 2;
    napi_value argv[2];
    CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    fs_render* r = get_render(env, argv[0]);
    if (!r) return NULL;
    double t = 5000;
    napi_get_value_double(env, argv[1], &t);
    int rc = fs_render_set_white_balance(r, t);
    if (rc) return throw_fs(env, rc);
    return NULL;
}
/* renderSetDepth(r, Float32Array | null, dw, dh) */
static napi_value js_render_set_depth(napi_env env, napi_callback_info info) {
    size_t argc = 4;
    napi_value argv[4];
    CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    fs_render* r = get_render(env, argv[0]);
    if (!r) return NULL;
    napi_valuetype t = napi_null;
    if (argc >= 2) napi_typeof(env, argv[1], &t);
    if (argc < 2 || t == napi_null || t == napi_undefined) {
        int rc = fs_render_set_depth(r, NULL, 0, 0);
        if (rc) return throw_fs(env, rc);
        return NULL;
    }
    napi_typedarray_type ty;
    size_t n = 0;
    void* data = NULL;
    if (napi_get_typedarray_info(env, argv[1], &ty, &n, &data, NULL, NULL) != napi_ok || ty != napi_float32_array) {
        napi_throw_type_error(env, NULL, "depth must be a Float32Array");
        return NULL;
    }
    double dw = 0, dh = 0;
    if (argc >= 4) { napi_get_value_double(env, argv[2], &dw); napi_get_value_double(env, argv[3], &dh); }
    if ((size_t)dw * (size_t)dh != n) {
        napi_throw_range_error(env, NULL, "depth length must be dw * dh");
        return NULL;
    }
    int rc = fs_render_set_depth(r, (const float*)data, (int)dw, (int)dh);
    if (rc) return throw_fs(env, rc);
    return NULL;
}
/* renderFrame(r, time, frame, jitX, jitY) */
static napi_value js_render_frame(napi_env env, napi_callback_info info) {
    size_t argc = 5;
    napi_value argv[5];
    CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    fs_render* r = get_render(env, argv[0]);
    if (!r) return NULL;
    double a[4] = {0, 0, 0, 0};
    for (size_t k = 1; k < argc && k < 5; k++) napi_get_value_double(env, argv[k], &a[k - 1]);
    int rc = fs_render_frame(r, a[0], a[1], (float)a[2], (float)a[3]);
    if (rc) return throw_fs(env, rc);
    return NULL;
}
/* renderSize(r) -> [vw, vh] */
static napi_value js_render_size(napi_env env, napi_callback_info info) {
    size_t argc = 1;
    napi_value argv[1];
    CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    fs_render* r = get_render(env, argv[0]);
    if (!r) return NULL;
    int vw = 0, vh = 0;
    int rc = fs_render_size(r, &vw, &vh);
    if (rc) return throw_fs(env, rc);
    napi_value arr, a, b;
    CHECK_NAPI(env, napi_create_array_with_length(env, 2, &arr));
    napi_create_int32(env, vw, &a);
    napi_create_int32(env, vh, &b);
    napi_set_element(env, arr, 0, a);
    napi_set_element(env, arr, 1, b);
    return arr;
}
/* renderRead(r, what, nFloats) -> Float32Array; renderUniforms(r) -> Float32Array(88) */
static napi_value js_render_read(napi_env env, napi_callback_info info) {
    size_t argc = 3;
    napi_value argv[3];
    CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    fs_render* r = get_render(env, argv[0]);
    if (!r) return NULL;
    double what = 0, nf = 0;
    napi_get_value_double(env, argv[1], &what);
    napi_get_value_double(env, argv[2], &nf);
    const size_t n = (size_t)nf;
    void* data = NULL;
    napi_value ab, arr;
    CHECK_NAPI(env, napi_create_arraybuffer(env, n * 4, &data, &ab));
    int rc = fs_render_read(r, (int)what, (float*)data, n * 4);
    if (rc) return throw_fs(env, rc);
    CHECK_NAPI(env, napi_create_typedarray(env, napi_float32_array, n, ab, 0, &arr));
    return arr;
}
static napi_value js_render_uniforms(napi_env env, napi_callback_info info) {
    size_t argc = 1;
    napi_value argv[1];
    CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    fs_render* r = get_render(env, argv[0]);
    if (!r) return NULL;
    float u[FS_RP_COUNT];
    int rc = fs_render_get_uniforms(r, u);
    if (rc) return throw_fs(env, rc);
    return float32_array(env, u, FS_RP_COUNT);
}
/* renderParamDefaults() -> Float32Array(88), renderParamIndex(name) */
static napi_value js_render_param_defaults(napi_env env, napi_callback_info info) {
    (void)info;
    float u[FS_RP_COUNT];
    fs_render_param_defaults(u);
    return float32_array(env, u, FS_RP_COUNT);
}
/* blackbodyLut(n, tmin, tmax, twb) -> Float32Array(2 n 4) (src/blackbody.js:78-88) */
static napi_value js_blackbody_lut(napi_env env, napi_callback_info info) {
    size_t argc = 4;
    napi_value argv[4];
    CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    double a[4] = {1024, 400, 3000, 5000};
    for (size_t k = 0; k < argc && k < 4; k++) napi_get_value_double(env, argv[k], &a[k]);
    const size_t n = (size_t)a[0] * 8;
    void* data = NULL;
    napi_value ab, arr;
    CHECK_NAPI(env, napi_create_arraybuffer(env, n * 4, &data, &ab));
    int rc = fs_blackbody_lut((int)a[0], a[1], a[2], a[3], (float*)data);
    if (rc) return throw_fs(env, rc);
    CHECK_NAPI(env, napi_create_typedarray(env, napi_float32_array, n, ab, 0, &arr));
    return arr;
}

static napi_value js_version(napi_env env, napi_callback_info info) {
    (void)info;
    napi_value s;
    napi_create_string_utf8(env, fs_version(), NAPI_AUTO_LENGTH, &s);
    return s;
}

#define EXPORT(name, fn)                                                     \
    do {                                                                     \
        napi_value f_;                                                       \
        napi_create_function(env, name, NAPI_AUTO_LENGTH, fn, NULL, &f_);    \
        napi_set_named_property(env, exports, name, f_);                     \
    } while (0)

static napi_value init(napi_env env, napi_value exports) {
    EXPORT("create", js_create);
    EXPORT("destroy", js_destroy);
    EXPORT("paramNames", js_param_names);
    EXPORT("paramDefaults", js_param_defaults);
    EXPORT("setParams", js_set_params);
    EXPORT("getParams", js_get_params);
    EXPORT("setLogs", js_set_logs);
    EXPORT("setSpots", js_set_spots);
    EXPORT("setFaceFlux", js_set_face_flux);
    EXPORT("reset", js_reset);
    EXPORT("step", js_step);
    EXPORT("stepAsync", js_step_async);
    EXPORT("sync", js_sync);
    EXPORT("readField", js_read_field);
    EXPORT("writeField", js_write_field);
    EXPORT("info", js_info);
    EXPORT("residual", js_residual);
    EXPORT("stats", js_stats);
    EXPORT("profileRows", js_profile_rows);
    EXPORT("diagEnable", js_diag_enable);
    EXPORT("diagRead", js_diag_read);
    EXPORT("profile", js_profile);
    EXPORT("timings", js_timings);
    EXPORT("buildCampfire", js_build_campfire);
    EXPORT("version", js_version);
    EXPORT("sparksCreate", js_sparks_create);
    EXPORT("sparksReset", js_sparks_reset);
    EXPORT("sparksSet", js_sparks_set);
    EXPORT("sparksBurst", js_sparks_burst);
    EXPORT("sparksStep", js_sparks_step);
    EXPORT("sparksRead", js_sparks_read);
    EXPORT("renderCreate", js_render_create);
    EXPORT("renderSet", js_render_set);
    EXPORT("renderSetCamera", js_render_set_camera);
    EXPORT("renderSetWhiteBalance", js_render_set_wb);
    EXPORT("renderSetDepth", js_render_set_depth);
    EXPORT("renderFrame", js_render_frame);
    EXPORT("renderSize", js_render_size);
    EXPORT("renderRead", js_render_read);
    EXPORT("renderUniforms", js_render_uniforms);
    EXPORT("renderParamDefaults", js_render_param_defaults);
    EXPORT("blackbodyLut", js_blackbody_lut);
    return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)
