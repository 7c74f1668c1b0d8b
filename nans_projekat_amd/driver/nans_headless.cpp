// nans_headless -- headless stand-in for the reference's platform loop.
//
// The reference's platform layer (code/sdl_nans.cpp:395-415, 913-1000) loads the
// simulation library with dlopen, binds its entry point with dlsym and calls it
// once per frame with the frame's dt.  This driver does the same with
// libnans_gpu.so and no window: it reads a scene file, advances it with the plain
// (drop-in) or the batched step, and writes a trajectory file for comparison and
// replay.
//
// Scene files are read by nans_scene_io.hpp (format described there).
//
// Trajectory file (binary, little endian):
//   char[8] "NANSTRJ1"; int32 nbodies, nframes, steps_per_frame;
//   then per frame: int32 step; float32 pos[nbodies][3]; float32 angles[nbodies][3]
// Frame 0 is the initial state.
//
// Usage: nans_headless --lib libnans_gpu.so --scene scene.txt --steps N
//                      [--every K] [--mode plain|batched] [--out traj.bin] [--parse-only]
//
// Platform-loop mode, the reference's frame loop (code/sdl_nans.cpp:541-555,
// 913-1000) around a SimUpdateAndRender library (libnans_sim.so, interface in
// include/nans_sim.h):
//        nans_headless --sim libnans_sim.so --frames N [--dt D] [--reload-every K]
//                      [--input script.txt] [--scene scene.txt] [--lib libnans_gpu.so]
//                      [--every K] [--out traj.bin]
// It maps the reference's 64 MiB PermanentStorage and 256 MiB TransientStorage,
// calls SimUpdateAndRender once per frame with a fixed dt (default 1/60), and
// unloads and reloads the library every K frames as the reference does when the
// library's file changes (code/sdl_nans.cpp:919-927).  --scene and --lib are
// passed on as NANS_SIM_SCENE and NANS_GPU_LIB.  The input script holds lines
//   <frame> <button> <0|1>          (button held from that frame on, or released)
//   <frame> mouse <xrel> <yrel>     (relative mouse motion in that frame)
// with buttons forward back left right shoot debug_up debug_down debug_left
// debug_right debug_forward debug_back debug_reset debug_continue.  Trajectory
// records are written after every K-th frame and after the last one; their step
// field is the number of frames run.
#include <dlfcn.h>
#include <sys/mman.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/nans_gpu.h"
#include "../../include/nans_sim.h"
#include "nans_scene_io.hpp"

using nans_io::planes;
using nans_io::scene;

struct nans_api {
    void *lib;
    void (*default_params)(nans_params *);
    const char *(*last_error)(void);
    int (*create)(nans_ctx **, const nans_params *);
    int (*destroy)(nans_ctx *);
    int (*upload)(nans_ctx *, int32_t, const int32_t *, const nans_bodies_soa *, const float *, const float *);
    int (*download)(nans_ctx *, nans_bodies_soa *);
    int (*step_plain)(nans_ctx *, int32_t);
    int (*step_batched)(nans_ctx *, int32_t);
};

static bool load_api(nans_api *a, const char *path)
{
    a->lib = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!a->lib) {
        fprintf(stderr, "nans_headless: cannot load %s: %s\n", path, dlerror());
        return false;
    }
    bool ok = true;
    auto bind = [&](void **slot, const char *sym) {
        *slot = dlsym(a->lib, sym);
        if (!*slot) {
            fprintf(stderr, "nans_headless: %s lacks %s\n", path, sym);
            ok = false;
        }
    };
    bind((void **)&a->default_params, "nans_default_params");
    bind((void **)&a->last_error, "nans_last_error");
    bind((void **)&a->create, "nans_create");
    bind((void **)&a->destroy, "nans_destroy");
    bind((void **)&a->upload, "nans_upload_worlds");
    bind((void **)&a->download, "nans_download_state");
    bind((void **)&a->step_plain, "nans_step_plain");
    bind((void **)&a->step_batched, "nans_step_batched");
    return ok;
}

static bool write_frame(FILE *f, int32_t step, const std::vector<float> &pos, const std::vector<float> &ang, size_t n)
{
    std::vector<float> buf(6 * n);
    for (size_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) {
            buf[3 * i + c] = pos[c * n + i];
            buf[3 * n + 3 * i + c] = ang[c * n + i];
        }
    return fwrite(&step, 4, 1, f) == 1 && fwrite(buf.data(), sizeof(float), buf.size(), f) == buf.size();
}


struct input_event {
    int frame, button;  // button -1: mouse
    int a, b;
};

static bool read_input(const char *path, std::vector<input_event> *ev)
{
    static const char *names[NANS_SIM_BUTTON_COUNT] = {
        "forward",    "back",        "left",          "right",      "shoot",       "debug_up",      "debug_down",
        "debug_left", "debug_right", "debug_forward", "debug_back", "debug_reset", "debug_continue"};
    FILE *f = fopen(path, "r");
    if (!f) {
        fprintf(stderr, "nans_headless: cannot open %s\n", path);
        return false;
    }
    char line[256], name[64];
    bool ok = true;
    while (ok && nans_io::next_line(f, line, sizeof(line))) {
        input_event e{0, -1, 0, 0};
        ok = sscanf(line, "%d %63s %d %d", &e.frame, name, &e.a, &e.b) >= 3 && e.frame >= 0;
        if (ok && strcmp(name, "mouse") != 0) {
            e.button = -2;
            for (int k = 0; k < NANS_SIM_BUTTON_COUNT; ++k)
                if (strcmp(name, names[k]) == 0) e.button = k;
            ok = e.button >= 0;
        }
        if (!ok) fprintf(stderr, "nans_headless: bad input line: %s", line);
        else ev->push_back(e);
    }
    fclose(f);
    return ok;
}

struct sim_code {
    void *lib;
    void (*update)(nans_sim_memory *, nans_sim_input *, void *, float);
    void (*release)(nans_sim_memory *);
};

// SDLLoadSimCode (code/sdl_nans.cpp:395-418)
static bool load_sim(sim_code *c, const char *path)
{
    c->lib = dlopen(path, RTLD_NOW | RTLD_GLOBAL);
    if (!c->lib) {
        fprintf(stderr, "nans_headless: cannot load %s: %s\n", path, dlerror());
        return false;
    }
    *(void **)&c->update = dlsym(c->lib, "SimUpdateAndRender");
    *(void **)&c->release = dlsym(c->lib, "nans_sim_release");
    if (!c->update) fprintf(stderr, "nans_headless: %s lacks SimUpdateAndRender\n", path);
    return c->update != nullptr;
}

static int run_sim(const char *simpath, const char *libpath, const char *scenepath, const char *inputpath,
                   const char *outpath, int frames, float dt, int reload_every, int every)
{
    if (libpath) setenv("NANS_GPU_LIB", libpath, 1);
    if (scenepath) setenv("NANS_SIM_SCENE", scenepath, 1);
    std::vector<input_event> events;
    if (inputpath && !read_input(inputpath, &events)) return 1;
    nans_sim_memory mem = {};
    mem.PermanentStorageSize = 64ull << 20;
    mem.TransientStorageSize = 256ull << 20;
    const size_t total = mem.PermanentStorageSize + mem.TransientStorageSize;
    void *base = mmap(nullptr, total, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);
    if (base == MAP_FAILED) {
        fprintf(stderr, "nans_headless: cannot map %zu bytes\n", total);
        return 1;
    }
    mem.PermanentStorage = base;
    mem.TransientStorage = (char *)base + mem.PermanentStorageSize;
    const nans_sim_state *st = (const nans_sim_state *)mem.PermanentStorage;
    sim_code code;
    if (!load_sim(&code, simpath)) return 1;
    nans_sim_input in = {};
    FILE *out = nullptr;
    const int per = every > 0 ? every : (frames > 0 ? frames : 1);
    int reloads = 0, rc = 0;
    auto t0 = std::chrono::steady_clock::now();
    for (int frame = 0; frame < frames && rc == 0; ++frame) {
        if (reload_every > 0 && frame > 0 && frame % reload_every == 0) {
            dlclose(code.lib);  // SDLUnloadSimCode (code/sdl_nans.cpp:420-431)
            if (!load_sim(&code, simpath)) return 1;
            ++reloads;
        }
        in.Sensitivity = 0.5f;  // code/sdl_nans.cpp:939
        in.XRel = in.YRel = 0;
        for (const input_event &e : events)
            if (e.frame == frame) {
                if (e.button < 0) in.XRel = e.a, in.YRel = e.b;
                else in.Buttons[e.button].EndedDown = e.a != 0;
            }
        code.update(&mem, &in, nullptr, dt);
        if (st->magic != NANS_SIM_MAGIC || st->status != NANS_OK) {
            fprintf(stderr, "nans_headless: frame %d: %s\n", frame, st->magic == NANS_SIM_MAGIC ? st->error :
                                                                     "the simulation did not initialise its state");
            rc = 1;
            break;
        }
        const int done = frame + 1;
        if (outpath && (done % per == 0 || done == frames)) {
            const int n = st->nboxes + st->nspheres;
            if (!out) {
                out = fopen(outpath, "wb");
                if (!out) {
                    fprintf(stderr, "nans_headless: cannot write %s\n", outpath);
                    rc = 1;
                    break;
                }
                const int32_t hdr[3] = {n, (frames + per - 1) / per, per};
                fwrite("NANSTRJ1", 1, 8, out);
                fwrite(hdr, 4, 3, out);
            }
            std::vector<float> P(st->pos, st->pos + 3 * n), A(st->angles, st->angles + 3 * n);
            if (!write_frame(out, done, P, A, (size_t)n)) rc = 1;
        }
    }
    const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (rc == 0)
        printf("sim: %d frames, %d bodies, %d reloads, %d uploads in %.3f s\n", st->frames,
               st->nboxes + st->nspheres, reloads, st->uploads, wall);
    if (out) fclose(out);
    if (code.release) code.release(&mem);
    dlclose(code.lib);
    munmap(base, total);
    return rc;
}

int main(int argc, char **argv)
{
    const char *libpath = nullptr, *scenepath = nullptr, *outpath = nullptr, *mode = "plain";
    const char *simpath = nullptr, *inputpath = nullptr;
    int steps = 0, every = 0, frames = 0, reload_every = 0;
    float dt = 1.0f / 60.0f;
    bool parse_only = false;
    for (int i = 1; i < argc; ++i) {
        auto arg = [&](const char *name) { return strcmp(argv[i], name) == 0 && i + 1 < argc; };
        if (arg("--lib")) libpath = argv[++i];
        else if (arg("--scene")) scenepath = argv[++i];
        else if (arg("--out")) outpath = argv[++i];
        else if (arg("--mode")) mode = argv[++i];
        else if (arg("--steps")) steps = atoi(argv[++i]);
        else if (arg("--every")) every = atoi(argv[++i]);
        else if (arg("--sim")) simpath = argv[++i];
        else if (arg("--input")) inputpath = argv[++i];
        else if (arg("--frames")) frames = atoi(argv[++i]);
        else if (arg("--reload-every")) reload_every = atoi(argv[++i]);
        else if (arg("--dt")) dt = strtof(argv[++i], nullptr);
        else if (strcmp(argv[i], "--parse-only") == 0) parse_only = true;
        else {
            fprintf(stderr, "nans_headless: unknown argument %s\n", argv[i]);
            return 2;
        }
    }
    if (simpath) {
        if (frames < 0 || every < 0 || reload_every < 0) {
            fprintf(stderr, "nans_headless: --frames, --every and --reload-every must be >= 0\n");
            return 2;
        }
        return run_sim(simpath, libpath, scenepath, inputpath, outpath, frames, dt, reload_every, every);
    }
    if (!scenepath || (!parse_only && !libpath) || steps < 0 || every < 0 ||
        (strcmp(mode, "plain") != 0 && strcmp(mode, "batched") != 0)) {
        fprintf(stderr, "usage: nans_headless --lib libnans_gpu.so --scene scene.txt --steps N [--every K] "
                        "[--mode plain|batched] [--out traj.bin] [--parse-only]\n"
                        "       nans_headless --sim libnans_sim.so --frames N [--dt D] [--reload-every K] "
                        "[--input script.txt] [--scene scene.txt] [--lib libnans_gpu.so] [--every K] [--out traj.bin]\n");
        return 2;
    }
    scene sc;
    memset(&sc.p, 0, sizeof(sc.p));
    if (!nans_io::read_scene(scenepath, &sc, "nans_headless")) return 1;
    const size_t n = sc.kind.size();
    const int nworlds = (int)sc.offsets.size() - 1;
    if (parse_only) {
        // a checksum of every value read, for round-trip tests without a GPU
        double sum = 0.0;
        for (auto *v : {&sc.pos, &sc.vel, &sc.angvel, &sc.angles, &sc.scale, &sc.mass, &sc.moi})
            for (float x : *v) sum += (double)x;
        printf("worlds %d bodies %zu iterations %d dt %.9g sum %.17g\n", nworlds, n, sc.p.solver_iterations,
               sc.p.dt, sum);
        return 0;
    }
    nans_api api;
    if (!load_api(&api, libpath)) return 1;
    nans_params p;
    api.default_params(&p);
    p.solver_iterations = sc.p.solver_iterations;
    p.beta = sc.p.beta;
    p.restitution = sc.p.restitution;
    p.friction = sc.p.friction;
    p.damping = sc.p.damping;
    p.gravity = sc.p.gravity;
    p.dt = sc.p.dt;
    p.max_gjk_iterations = sc.p.max_gjk_iterations;
    p.max_epa_iterations = sc.p.max_epa_iterations;
    p.epa_tolerance = sc.p.epa_tolerance;
    p.fp64_islands = sc.p.fp64_islands;
    memcpy(p.floor_pos, sc.p.floor_pos, sizeof(p.floor_pos));
    memcpy(p.floor_scale, sc.p.floor_scale, sizeof(p.floor_scale));
    p.floor_mass = sc.p.floor_mass;
    p.floor_moi = sc.p.floor_moi;
    // per-world coefficients: NaN entries take the params' values
    std::vector<float> fric(sc.fric), rest(sc.rest);
    bool any_f = false, any_r = false;
    for (int w = 0; w < nworlds; ++w) {
        if (fric[w] != fric[w]) fric[w] = p.friction;
        else any_f = true;
        if (rest[w] != rest[w]) rest[w] = p.restitution;
        else any_r = true;
    }
    std::vector<float> P = planes(sc.pos, n), V = planes(sc.vel, n), W = planes(sc.angvel, n),
                       A = planes(sc.angles, n), S = planes(sc.scale, n);
    nans_bodies_soa b = {P.data(), V.data(), W.data(), A.data(), nullptr, nullptr, S.data(), sc.mass.data(),
                         sc.moi.data(), sc.kind.data()};
    nans_ctx *ctx = nullptr;
    int rc = api.create(&ctx, &p);
    if (rc == NANS_OK)
        rc = api.upload(ctx, nworlds, sc.offsets.data(), &b, any_f ? fric.data() : nullptr,
                        any_r ? rest.data() : nullptr);
    FILE *out = nullptr;
    const int per = every > 0 ? every : (steps > 0 ? steps : 1);
    const int nframes = 1 + (steps + per - 1) / per;
    if (rc == NANS_OK && outpath) {
        out = fopen(outpath, "wb");
        if (!out) {
            fprintf(stderr, "nans_headless: cannot write %s\n", outpath);
            rc = NANS_ERR_INVALID;
        } else {
            const int32_t hdr[3] = {(int32_t)n, nframes, per};
            fwrite("NANSTRJ1", 1, 8, out);
            fwrite(hdr, 4, 3, out);
            write_frame(out, 0, P, A, n);
        }
    }
    auto t0 = std::chrono::steady_clock::now();
    double busy = 0.0;
    for (int done = 0; rc == NANS_OK && done < steps;) {
        const int k = per < steps - done ? per : steps - done;
        auto s0 = std::chrono::steady_clock::now();
        rc = strcmp(mode, "plain") == 0 ? api.step_plain(ctx, k) : api.step_batched(ctx, k);
        busy += std::chrono::duration<double>(std::chrono::steady_clock::now() - s0).count();
        done += k;
        if (rc == NANS_OK && out) {
            nans_bodies_soa d = {P.data(), V.data(), W.data(), A.data(), nullptr, nullptr, nullptr, nullptr, nullptr,
                                 nullptr};
            rc = api.download(ctx, &d);
            if (rc == NANS_OK && !write_frame(out, done, P, A, n)) rc = NANS_ERR_INVALID;
        }
    }
    const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (rc != NANS_OK) fprintf(stderr, "nans_headless: error %d: %s\n", rc, api.last_error());
    else
        printf("%s: %d worlds, %zu bodies, %d steps in %.3f s (stepping %.3f s): %.1f body-steps/s\n", mode,
               nworlds, n, steps, wall, busy, busy > 0 ? (double)n * steps / busy : 0.0);
    if (out) fclose(out);
    if (ctx) api.destroy(ctx);
    return rc == NANS_OK ? 0 : 1;
}
