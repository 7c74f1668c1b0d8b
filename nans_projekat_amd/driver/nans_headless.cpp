// nans_headless -- headless stand-in for the reference's platform loop.
//
// The reference's platform layer (code/sdl_nans.cpp:395-415, 913-1000) loads the
// simulation library with dlopen, binds its entry point with dlsym and calls it
// once per frame with the frame's dt.  This driver does the same with
// libnans_gpu.so and no window: it reads a scene file, advances it with the plain
// (drop-in) or the batched step, and writes a trajectory file for comparison and
// replay.
//
// Scene file (text, '#' starts a comment line; floats as written by "%.9g" read
// back exactly):
//   nans-scene 1
//   params <solver_iterations> <beta> <restitution> <friction> <damping> <gravity> <dt>
//          <max_gjk_iterations> <max_epa_iterations> <epa_tolerance> <fp64_islands>
//          <floor_pos x y z> <floor_scale x y z> <floor_mass> <floor_moi>      (one line)
//   worlds <nworlds>
//   world <nbodies> <friction|-> <restitution|->      (then nbodies body lines)
//   body <box|sphere> <px py pz> <vx vy vz> <wx wy wz> <ax ay az> <sx sy sz> <mass> <moi>
// Boxes precede spheres within a world; for a sphere sx is the radius.
//
// Trajectory file (binary, little endian):
//   char[8] "NANSTRJ1"; int32 nbodies, nframes, steps_per_frame;
//   then per frame: int32 step; float32 pos[nbodies][3]; float32 angles[nbodies][3]
// Frame 0 is the initial state.
//
// Usage: nans_headless --lib libnans_gpu.so --scene scene.txt --steps N
//                      [--every K] [--mode plain|batched] [--out traj.bin] [--parse-only]
#include <dlfcn.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/nans_gpu.h"

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

struct scene {
    nans_params p;
    std::vector<int32_t> offsets;
    std::vector<float> fric, rest;     // per world (NaN: the params' value)
    std::vector<int32_t> kind;
    std::vector<float> pos, vel, angvel, angles, scale, mass, moi;  // [N][3] / [N] as read
};

static bool next_line(FILE *f, char *buf, int n)
{
    while (fgets(buf, n, f)) {
        char *s = buf;
        while (*s == ' ' || *s == '\t') ++s;
        if (*s == '#' || *s == '\n' || *s == '\r' || *s == 0) continue;
        return true;
    }
    return false;
}

static bool parse_coef(const char *tok, float *out)
{
    if (strcmp(tok, "-") == 0) {
        *out = NAN;
        return true;
    }
    char *end = nullptr;
    *out = strtof(tok, &end);
    return end && *end == 0;
}

// Reads a scene; p starts from the library defaults (or zeros with --parse-only).
static bool read_scene(const char *path, scene *sc)
{
    FILE *f = fopen(path, "r");
    if (!f) {
        fprintf(stderr, "nans_headless: cannot open %s\n", path);
        return false;
    }
    char line[1024];
    bool ok = next_line(f, line, sizeof(line)) && strncmp(line, "nans-scene 1", 12) == 0;
    if (!ok) fprintf(stderr, "nans_headless: %s is not a nans-scene 1 file\n", path);
    nans_params &p = sc->p;
    if (ok) {
        ok = next_line(f, line, sizeof(line)) &&
             sscanf(line, "params %d %f %f %f %f %f %f %d %d %f %d %f %f %f %f %f %f %f %f", &p.solver_iterations,
                    &p.beta, &p.restitution, &p.friction, &p.damping, &p.gravity, &p.dt, &p.max_gjk_iterations,
                    &p.max_epa_iterations, &p.epa_tolerance, &p.fp64_islands, &p.floor_pos[0], &p.floor_pos[1],
                    &p.floor_pos[2], &p.floor_scale[0], &p.floor_scale[1], &p.floor_scale[2], &p.floor_mass,
                    &p.floor_moi) == 19;
        if (!ok) fprintf(stderr, "nans_headless: bad params line\n");
    }
    int nworlds = 0;
    if (ok) {
        ok = next_line(f, line, sizeof(line)) && sscanf(line, "worlds %d", &nworlds) == 1 && nworlds >= 1;
        if (!ok) fprintf(stderr, "nans_headless: bad worlds line\n");
    }
    sc->offsets.assign(1, 0);
    for (int w = 0; ok && w < nworlds; ++w) {
        int nb = 0;
        char fr[64], rs[64];
        ok = next_line(f, line, sizeof(line)) && sscanf(line, "world %d %63s %63s", &nb, fr, rs) == 3 && nb >= 0;
        float cf = 0.0f, cr = 0.0f;
        ok = ok && parse_coef(fr, &cf) && parse_coef(rs, &cr);
        if (!ok) {
            fprintf(stderr, "nans_headless: bad world line %d\n", w);
            break;
        }
        sc->fric.push_back(cf);
        sc->rest.push_back(cr);
        for (int i = 0; ok && i < nb; ++i) {
            char kind[16];
            float v[17];
            ok = next_line(f, line, sizeof(line)) &&
                 sscanf(line, "body %15s %f %f %f %f %f %f %f %f %f %f %f %f %f %f %f %f %f", kind, &v[0], &v[1],
                        &v[2], &v[3], &v[4], &v[5], &v[6], &v[7], &v[8], &v[9], &v[10], &v[11], &v[12], &v[13],
                        &v[14], &v[15], &v[16]) == 18 &&
                 (strcmp(kind, "box") == 0 || strcmp(kind, "sphere") == 0);
            if (!ok) {
                fprintf(stderr, "nans_headless: bad body line (world %d, body %d)\n", w, i);
                break;
            }
            sc->kind.push_back(strcmp(kind, "box") == 0 ? NANS_SHAPE_BOX : NANS_SHAPE_SPHERE);
            for (int c = 0; c < 3; ++c) {
                sc->pos.push_back(v[c]);
                sc->vel.push_back(v[3 + c]);
                sc->angvel.push_back(v[6 + c]);
                sc->angles.push_back(v[9 + c]);
                sc->scale.push_back(v[12 + c]);
            }
            sc->mass.push_back(v[15]);
            sc->moi.push_back(v[16]);
        }
        sc->offsets.push_back((int32_t)sc->kind.size());
    }
    fclose(f);
    return ok;
}

// [N][3] -> [3][N] planes
static std::vector<float> planes(const std::vector<float> &a, size_t n)
{
    std::vector<float> out(3 * n);
    for (size_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) out[c * n + i] = a[3 * i + c];
    return out;
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

int main(int argc, char **argv)
{
    const char *libpath = nullptr, *scenepath = nullptr, *outpath = nullptr, *mode = "plain";
    int steps = 0, every = 0;
    bool parse_only = false;
    for (int i = 1; i < argc; ++i) {
        auto arg = [&](const char *name) { return strcmp(argv[i], name) == 0 && i + 1 < argc; };
        if (arg("--lib")) libpath = argv[++i];
        else if (arg("--scene")) scenepath = argv[++i];
        else if (arg("--out")) outpath = argv[++i];
        else if (arg("--mode")) mode = argv[++i];
        else if (arg("--steps")) steps = atoi(argv[++i]);
        else if (arg("--every")) every = atoi(argv[++i]);
        else if (strcmp(argv[i], "--parse-only") == 0) parse_only = true;
        else {
            fprintf(stderr, "nans_headless: unknown argument %s\n", argv[i]);
            return 2;
        }
    }
    if (!scenepath || (!parse_only && !libpath) || steps < 0 || every < 0 ||
        (strcmp(mode, "plain") != 0 && strcmp(mode, "batched") != 0)) {
        fprintf(stderr, "usage: nans_headless --lib libnans_gpu.so --scene scene.txt --steps N [--every K] "
                        "[--mode plain|batched] [--out traj.bin] [--parse-only]\n");
        return 2;
    }
    scene sc;
    memset(&sc.p, 0, sizeof(sc.p));
    if (!read_scene(scenepath, &sc)) return 1;
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
