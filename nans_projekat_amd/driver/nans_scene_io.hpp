// nans_scene_io.hpp -- reader of the text scene format shared by the headless
// driver (nans_headless.cpp) and the SimUpdateAndRender adapter (nans_sim.cpp).
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
#ifndef NANS_SCENE_IO_HPP
#define NANS_SCENE_IO_HPP

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/nans_gpu.h"

namespace nans_io {

struct scene {
    nans_params p;
    std::vector<int32_t> offsets;
    std::vector<float> fric, rest;     // per world (NaN: the params' value)
    std::vector<int32_t> kind;
    std::vector<float> pos, vel, angvel, angles, scale, mass, moi;  // [N][3] / [N] as read
};

inline bool next_line(FILE *f, char *buf, int n)
{
    while (fgets(buf, n, f)) {
        char *s = buf;
        while (*s == ' ' || *s == '\t') ++s;
        if (*s == '#' || *s == '\n' || *s == '\r' || *s == 0) continue;
        return true;
    }
    return false;
}

inline bool parse_coef(const char *tok, float *out)
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
inline bool read_scene(const char *path, scene *sc, const char *who)
{
    FILE *f = fopen(path, "r");
    if (!f) {
        fprintf(stderr, "%s: cannot open %s\n", who, path);
        return false;
    }
    char line[1024];
    bool ok = next_line(f, line, sizeof(line)) && strncmp(line, "nans-scene 1", 12) == 0;
    if (!ok) fprintf(stderr, "%s: %s is not a nans-scene 1 file\n", who, path);
    nans_params &p = sc->p;
    if (ok) {
        ok = next_line(f, line, sizeof(line)) &&
             sscanf(line, "params %d %f %f %f %f %f %f %d %d %f %d %f %f %f %f %f %f %f %f", &p.solver_iterations,
                    &p.beta, &p.restitution, &p.friction, &p.damping, &p.gravity, &p.dt, &p.max_gjk_iterations,
                    &p.max_epa_iterations, &p.epa_tolerance, &p.fp64_islands, &p.floor_pos[0], &p.floor_pos[1],
                    &p.floor_pos[2], &p.floor_scale[0], &p.floor_scale[1], &p.floor_scale[2], &p.floor_mass,
                    &p.floor_moi) == 19;
        if (!ok) fprintf(stderr, "%s: bad params line\n", who);
    }
    int nworlds = 0;
    if (ok) {
        ok = next_line(f, line, sizeof(line)) && sscanf(line, "worlds %d", &nworlds) == 1 && nworlds >= 1;
        if (!ok) fprintf(stderr, "%s: bad worlds line\n", who);
    }
    sc->offsets.assign(1, 0);
    for (int w = 0; ok && w < nworlds; ++w) {
        int nb = 0;
        char fr[64], rs[64];
        ok = next_line(f, line, sizeof(line)) && sscanf(line, "world %d %63s %63s", &nb, fr, rs) == 3 && nb >= 0;
        float cf = 0.0f, cr = 0.0f;
        ok = ok && parse_coef(fr, &cf) && parse_coef(rs, &cr);
        if (!ok) {
            fprintf(stderr, "%s: bad world line %d\n", who, w);
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
                fprintf(stderr, "%s: bad body line (world %d, body %d)\n", who, w, i);
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
inline std::vector<float> planes(const std::vector<float> &a, size_t n)
{
    std::vector<float> out(3 * n);
    for (size_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) out[c * n + i] = a[3 * i + c];
    return out;
}

}  // namespace nans_io

#endif
