// nans_sim.cpp -> libnans_sim.so: SimUpdateAndRender with the physics block on
// the GPU (interface and storage layout in include/nans_sim.h).
//
// Per frame, in the reference's order (code/nans.cpp:1719-1772): Init on the
// first call, HandleInput, UpdateCamera, the physics block, the camera's Front
// from Yaw/Pitch.  The physics block is libnans_gpu.so's plain step on the
// world kept on the device; the host copy in PermanentStorage is refreshed after
// every frame and written back to the device only when input has edited it.
//
// One difference from the reference, shared with every caller of the library:
// a body moved by input between frames collides with its corners at the new
// pose in that frame's narrowphase (the step rebuilds each box's corners from
// its pose), while the reference reads the Model its render built in the
// previous frame (code/nans.cpp:979-981, 1913-1924) and so sees the move one
// frame later.
#include <dlfcn.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/nans_sim.h"
#include "nans_scene_io.hpp"

namespace {

struct gpu_api {
    bool bound;
    void (*default_params)(nans_params *);
    const char *(*last_error)(void);
    int (*create)(nans_ctx **, const nans_params *);
    int (*destroy)(nans_ctx *);
    int (*set_params)(nans_ctx *, const nans_params *);
    int (*upload)(nans_ctx *, int32_t, const int32_t *, const nans_bodies_soa *, const float *, const float *);
    int (*set_state)(nans_ctx *, const nans_bodies_soa *);
    int (*download)(nans_ctx *, nans_bodies_soa *);
    int (*step_plain)(nans_ctx *, int32_t);
};

// Bound once per loaded image of this library; a reload binds again.
gpu_api g_api;

void fail(nans_sim_state *st, int rc, const char *fmt, ...)
{
    st->status = rc;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(st->error, sizeof(st->error), fmt, ap);
    va_end(ap);
    fprintf(stderr, "nans_sim: %s\n", st->error);
}

// libnans_gpu.so: $NANS_GPU_LIB, else the file beside this library.
std::string gpu_lib_path()
{
    const char *env = getenv("NANS_GPU_LIB");
    if (env && *env) return env;
    Dl_info info;
    if (dladdr((void *)&gpu_lib_path, &info) && info.dli_fname) {
        std::string self = info.dli_fname;
        size_t slash = self.rfind('/');
        return (slash == std::string::npos ? std::string(".") : self.substr(0, slash)) + "/libnans_gpu.so";
    }
    return "libnans_gpu.so";
}

bool bind_gpu(nans_sim_state *st)
{
    if (g_api.bound) return true;
    const std::string path = gpu_lib_path();
    // RTLD_NODELETE: the HIP runtime and the context's device memory outlive
    // reloads of this library.
    void *lib = dlopen(path.c_str(), RTLD_NOW | RTLD_LOCAL | RTLD_NODELETE);
    if (!lib) {
        fail(st, NANS_ERR_STATE, "cannot load %s: %s", path.c_str(), dlerror());
        return false;
    }
    struct {
        void **slot;
        const char *name;
    } syms[] = {{(void **)&g_api.default_params, "nans_default_params"},
                {(void **)&g_api.last_error, "nans_last_error"},
                {(void **)&g_api.create, "nans_create"},
                {(void **)&g_api.destroy, "nans_destroy"},
                {(void **)&g_api.set_params, "nans_set_params"},
                {(void **)&g_api.upload, "nans_upload_worlds"},
                {(void **)&g_api.set_state, "nans_set_state"},
                {(void **)&g_api.download, "nans_download_state"},
                {(void **)&g_api.step_plain, "nans_step_plain"}};
    for (auto &s : syms) {
        *s.slot = dlsym(lib, s.name);
        if (!*s.slot) {
            fail(st, NANS_ERR_STATE, "%s lacks %s", path.c_str(), s.name);
            return false;
        }
    }
    g_api.bound = true;
    return true;
}

int nbodies(const nans_sim_state *st) { return st->nboxes + st->nspheres; }

nans_bodies_soa soa(nans_sim_state *st)
{
    return {st->pos, st->vel, st->angvel, st->angles, st->force, st->torque, st->scale, st->mass, st->moi, st->kind};
}

// glm's vec3 arithmetic in binary32: dot sums (x*x + y*y) + z*z, normalize
// multiplies by 1/sqrt(dot).
struct v3 {
    float x, y, z;
};
v3 load(const float *a) { return {a[0], a[1], a[2]}; }
void store(float *a, v3 v) { a[0] = v.x, a[1] = v.y, a[2] = v.z; }
v3 cross(v3 a, v3 b) { return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
v3 normalize(v3 v)
{
    const float d = (v.x * v.x + v.y * v.y) + v.z * v.z;
    const float r = 1.0f / sqrtf(d);
    return {v.x * r, v.y * r, v.z * r};
}
float radians(float deg) { return deg * (float)0.01745329251994329576923690768489; }

// Body i of the [3][n] planes.
v3 get(const float *a, int n, int i) { return {a[i], a[n + i], a[2 * n + i]}; }
void put(float *a, int n, int i, v3 v) { a[i] = v.x, a[n + i] = v.y, a[2 * n + i] = v.z; }

// The physics part of the reference's Init (code/nans.cpp:1597-1678): three unit
// cubes, the box Cubes[3] drawn with Model scale (0.5, 1, 0.5) (code/nans.cpp:
// 1938-1940) and keeping the unit-cube MOI, one sphere.  MOI as the reference
// computes it, with pow in binary64.
void demo_scene(nans_sim_state *st)
{
    st->nboxes = 4;
    st->nspheres = 1;
    const int n = 5;
    const float bp[4][3] = {{2.0f, 3.5f, 2.0f}, {2.0f, 1.0f, 2.0f}, {2.0f, 4.5f, 2.0f}, {1.0f, 1.0f, 1.0f}};
    const float bs[4][3] = {{1.0f, 1.0f, 1.0f}, {1.0f, 1.0f, 1.0f}, {1.0f, 1.0f, 1.0f}, {0.5f, 1.0f, 0.5f}};
    for (int i = 0; i < 4; ++i) {
        put(st->pos, n, i, {bp[i][0], bp[i][1], bp[i][2]});
        put(st->scale, n, i, {bs[i][0], bs[i][1], bs[i][2]});
        st->mass[i] = 1.0f;
        const float size = 1.0f;
        st->moi[i] = (float)((st->mass[i] / 12.0f) * (2.0f * pow(size, 2)));
        st->kind[i] = NANS_SHAPE_BOX;
    }
    put(st->pos, n, 4, {0.1f, 1.1f, 1.1f});
    put(st->scale, n, 4, {0.25f, 0.0f, 0.0f});
    st->mass[4] = 2.0f;
    st->moi[4] = (float)((2.0f / 5.0f) * st->mass[4] * pow(0.25f, 2));
    st->kind[4] = NANS_SHAPE_SPHERE;
    for (float *a : {st->vel, st->angvel, st->angles, st->force, st->torque})
        for (int k = 0; k < 3 * n; ++k) a[k] = 0.0f;
}

bool load_scene(nans_sim_state *st, const char *path)
{
    nans_io::scene sc;
    sc.p = st->params;
    if (!nans_io::read_scene(path, &sc, "nans_sim")) {
        fail(st, NANS_ERR_INVALID, "cannot read scene %s", path);
        return false;
    }
    const int n = (int)sc.kind.size();
    if (sc.offsets.size() != 2) {
        fail(st, NANS_ERR_INVALID, "scene %s holds %d worlds; the simulation runs one", path,
             (int)sc.offsets.size() - 1);
        return false;
    }
    if (n > NANS_SIM_MAX_BODIES) {
        fail(st, NANS_ERR_CAPACITY, "scene %s holds %d bodies (at most %d)", path, n, NANS_SIM_MAX_BODIES);
        return false;
    }
    int nb = 0;
    while (nb < n && sc.kind[nb] == NANS_SHAPE_BOX) ++nb;
    for (int i = nb; i < n; ++i)
        if (sc.kind[i] != NANS_SHAPE_SPHERE) {
            fail(st, NANS_ERR_INVALID, "scene %s: boxes must precede spheres", path);
            return false;
        }
    st->params = sc.p;
    if (sc.fric[0] == sc.fric[0]) st->params.friction = sc.fric[0];
    if (sc.rest[0] == sc.rest[0]) st->params.restitution = sc.rest[0];
    st->nboxes = nb;
    st->nspheres = n - nb;
    const std::vector<float> *src[5] = {&sc.pos, &sc.vel, &sc.angvel, &sc.angles, &sc.scale};
    float *dst[5] = {st->pos, st->vel, st->angvel, st->angles, st->scale};
    for (int f = 0; f < 5; ++f)
        for (int i = 0; i < n; ++i)
            for (int c = 0; c < 3; ++c) dst[f][c * n + i] = (*src[f])[3 * i + c];
    for (int i = 0; i < n; ++i) {
        st->mass[i] = sc.mass[i];
        st->moi[i] = sc.moi[i];
        st->kind[i] = sc.kind[i];
    }
    for (int k = 0; k < 3 * n; ++k) st->force[k] = st->torque[k] = 0.0f;
    return true;
}

// Init (code/nans.cpp:1585-1595 camera, 1597-1678 bodies); the device context
// is created on the first frame.
void init(nans_sim_state *st)
{
    g_api.default_params(&st->params);
    nans_sim_camera &c = st->camera;
    c.pitch = 0.0f;
    c.yaw = -90.0f;
    c.speed = 0.05f;
    store(c.position, {0.0f, 0.0f, 3.0f});
    const v3 up = {0.0f, 1.0f, 0.0f}, target = {0.0f, 0.0f, 0.0f};
    const v3 p = load(c.position);
    const v3 dir = normalize({p.x - target.x, p.y - target.y, p.z - target.z});
    store(c.front, {0.0f, 0.0f, -1.0f});
    const v3 right = normalize(cross(up, dir));
    store(c.up, cross(dir, right));
    const char *scene = getenv("NANS_SIM_SCENE");
    if (scene && *scene) load_scene(st, scene);
    else demo_scene(st);
}

// HandleInput (code/nans.cpp:122-185) and ShootSphere (:112-120).  Returns
// whether a body was edited.
bool handle_input(nans_sim_state *st, const nans_sim_input *in, float dt)
{
    nans_sim_camera &c = st->camera;
    const int n = nbodies(st);
    bool edited = false;
    auto down = [&](int b) { return in->Buttons[b].EndedDown != 0; };
    v3 p = load(c.position);
    const v3 f = load(c.front);
    if (down(NANS_SIM_MOVE_FORWARD)) p = {p.x + c.speed * f.x, p.y + c.speed * f.y, p.z + c.speed * f.z};
    if (down(NANS_SIM_MOVE_LEFT)) {
        const v3 r = normalize(cross(f, load(c.up)));
        p = {p.x - r.x * c.speed, p.y - r.y * c.speed, p.z - r.z * c.speed};
    }
    if (down(NANS_SIM_MOVE_BACK)) p = {p.x - c.speed * f.x, p.y - c.speed * f.y, p.z - c.speed * f.z};
    if (down(NANS_SIM_MOVE_RIGHT)) {
        const v3 r = normalize(cross(f, load(c.up)));
        p = {p.x + r.x * c.speed, p.y + r.y * c.speed, p.z + r.z * c.speed};
    }
    store(c.position, p);
    if (down(NANS_SIM_SHOOT) && st->nspheres > 0) {
        const int s = st->nboxes;  // Spheres[0]
        const float shoot = 4001.0f;  // SHOOT_FORCE (code/nans.h:48)
        put(st->force, n, s, {0.0f + shoot * f.x, 0.0f + shoot * f.y, 0.0f + shoot * f.z});
        put(st->torque, n, s, {0.0f, 0.0f, 0.0f});
        put(st->vel, n, s, {0.0f, 0.0f, 0.0f});
        put(st->angvel, n, s, {0.0f, 0.0f, 0.0f});
        put(st->pos, n, s, p);
        edited = true;
    }
    // Cubes[3] nudged by dt along an axis (code/nans.cpp:146-170)
    const struct {
        int button;
        v3 axis;
    } nudges[] = {{NANS_SIM_DEBUG_LEFT, {1.0f, 0.0f, 0.0f}},  {NANS_SIM_DEBUG_RIGHT, {-1.0f, 0.0f, 0.0f}},
                  {NANS_SIM_DEBUG_UP, {0.0f, 1.0f, 0.0f}},    {NANS_SIM_DEBUG_DOWN, {0.0f, -1.0f, 0.0f}},
                  {NANS_SIM_DEBUG_FORWARD, {0.0f, 0.0f, 1.0f}}, {NANS_SIM_DEBUG_BACK, {0.0f, 0.0f, -1.0f}}};
    for (const auto &u : nudges)
        if (down(u.button) && st->nboxes > 3) {
            const v3 q = get(st->pos, n, 3);
            put(st->pos, n, 3, {q.x + dt * u.axis.x, q.y + dt * u.axis.y, q.z + dt * u.axis.z});
            edited = true;
        }
    if (down(NANS_SIM_DEBUG_RESET)) {  // code/nans.cpp:171-184
        const v3 reset[4] = {{2.0f, 3.5f, 2.0f}, {2.0f, 1.0f, 2.0f}, {2.0f, 4.5f, 2.0f}, {0.0f, 0.0f, 0.0f}};
        for (int i = 0; i < 4 && i < st->nboxes; ++i) {
            put(st->pos, n, i, reset[i]);
            put(st->vel, n, i, {0.0f, 0.0f, 0.0f});
            put(st->angvel, n, i, {0.0f, 0.0f, 0.0f});
            edited = true;
        }
    }
    return edited;
}

// UpdateCamera (code/nans.cpp:19-36); like the reference it consumes the
// mouse's relative motion.
void update_camera(nans_sim_state *st, nans_sim_input *in)
{
    nans_sim_camera &c = st->camera;
    c.yaw += in->XRel * in->Sensitivity;
    c.pitch += -in->YRel * in->Sensitivity;
    in->XRel = 0;
    in->YRel = 0;
    if (c.pitch > 89.0f) c.pitch = 89.0f;
    if (c.pitch < -89.0f) c.pitch = -89.0f;
    if (c.position[1] < 0.5) c.position[1] = 0.5;
}

// The camera's Front from Yaw and Pitch (code/nans.cpp:1768-1772).
void update_front(nans_sim_state *st)
{
    nans_sim_camera &c = st->camera;
    v3 f;
    f.x = cosf(radians(c.yaw)) * cosf(radians(c.pitch));
    f.y = sinf(radians(c.pitch));
    f.z = sinf(radians(c.yaw)) * cosf(radians(c.pitch));
    store(c.front, normalize(f));
}

bool gpu_ok(nans_sim_state *st, int rc, const char *what)
{
    if (rc == NANS_OK) return true;
    fail(st, rc, "%s: %s", what, g_api.last_error());
    return false;
}

}  // namespace

extern "C" void SimUpdateAndRender(nans_sim_memory *Memory, nans_sim_input *Input, void *Render, float dt)
{
    (void)Render;  // nothing is drawn
    if (!Memory || !Memory->PermanentStorage) return;
    nans_sim_state *st = (nans_sim_state *)Memory->PermanentStorage;
    if (Memory->PermanentStorageSize < sizeof(nans_sim_state)) {
        if (!Memory->IsInitialized) {
            fprintf(stderr, "nans_sim: PermanentStorage holds %llu bytes, the state needs %zu\n",
                    (unsigned long long)Memory->PermanentStorageSize, sizeof(nans_sim_state));
            Memory->IsInitialized = 1;
        }
        return;
    }
    if (!Memory->IsInitialized) {
        memset(st, 0, sizeof(*st));
        st->magic = NANS_SIM_MAGIC;
        if (bind_gpu(st)) init(st);
        Memory->IsInitialized = 1;
    }
    if (st->magic != NANS_SIM_MAGIC || st->status != NANS_OK) return;
    if (!bind_gpu(st)) return;
    const int n = nbodies(st);
    nans_bodies_soa b = soa(st);
    if (!st->gpu) {
        if (!gpu_ok(st, g_api.create(&st->gpu, &st->params), "nans_create")) return;
        const int32_t offsets[2] = {0, n};
        if (!gpu_ok(st, g_api.upload(st->gpu, 1, offsets, &b, nullptr, nullptr), "nans_upload_worlds")) return;
        ++st->uploads;
    }
    nans_sim_input none;
    memset(&none, 0, sizeof(none));
    nans_sim_input *in = Input ? Input : &none;
    const bool edited = handle_input(st, in, dt);
    update_camera(st, in);
    if (edited && !gpu_ok(st, g_api.set_state(st->gpu, &b), "nans_set_state")) return;
    if (dt != st->params.dt || st->frames == 0) {
        st->params.dt = dt;
        if (!gpu_ok(st, g_api.set_params(st->gpu, &st->params), "nans_set_params")) return;
    }
    if (!gpu_ok(st, g_api.step_plain(st->gpu, 1), "nans_step_plain")) return;
    nans_bodies_soa d = {st->pos, st->vel, st->angvel, st->angles, st->force, st->torque,
                         nullptr, nullptr, nullptr, nullptr};
    if (!gpu_ok(st, g_api.download(st->gpu, &d), "nans_download_state")) return;
    ++st->frames;
    update_front(st);
}

extern "C" void nans_sim_release(nans_sim_memory *Memory)
{
    if (!Memory || !Memory->PermanentStorage || Memory->PermanentStorageSize < sizeof(nans_sim_state)) return;
    nans_sim_state *st = (nans_sim_state *)Memory->PermanentStorage;
    if (st->magic == NANS_SIM_MAGIC && st->gpu && bind_gpu(st)) g_api.destroy(st->gpu);
    st->gpu = nullptr;
    Memory->IsInitialized = 0;
}
