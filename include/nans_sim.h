/*
 * nans_sim.h -- the plugin interface of libnans_sim.so, a SimUpdateAndRender
 * replacement that runs the physics block on the GPU.
 *
 * The reference's platform layer (code/sdl_nans.cpp:395-431, 541-555, 913-986)
 * loads its simulation library with dlopen, binds SimUpdateAndRender with dlsym,
 * calls it once per frame with the platform's memory block, the frame's input
 * and dt, and reloads the library whenever its file changes.  libnans_sim.so
 * exports the same symbol with the same argument layout, so that loop can load
 * it in place of nans.so.  It runs the reference's per-frame work that touches
 * the simulation: Init on the first call (code/nans.cpp:1724-1729), the input
 * that moves bodies or the camera (HandleInput / UpdateCamera, code/nans.cpp:
 * 19-36, 112-185), the physics block (code/nans.cpp:1756-1762) through
 * libnans_gpu.so's plain step, and the camera update (code/nans.cpp:1766-1772).
 * It draws nothing and leaves Render untouched.
 *
 * Everything the simulation keeps between frames lives in PermanentStorage as a
 * nans_sim_state, including the device context: a reloaded libnans_sim.so picks
 * the world up where the previous image left it, without a new upload.
 * libnans_gpu.so is opened with RTLD_NODELETE, so it (and the device memory the
 * context owns) stays loaded across reloads.
 */
#ifndef NANS_SIM_H
#define NANS_SIM_H

#include <stdint.h>

#include "nans_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* memory (code/nans.h:227-235), the platform's storage block. */
typedef struct nans_sim_memory {
    void *PermanentStorage;
    uint64_t PermanentStorageSize;
    void *TransientStorage;
    uint64_t TransientStorageSize;
    int32_t IsInitialized;
} nans_sim_memory;

/* sdl_button_state, sdl_keyboard, sdl_mouse, sdl_input (code/nans.h:119-165). */
typedef struct nans_sim_button {
    int32_t HalfTransitionCount;
    int32_t EndedDown;
} nans_sim_button;

enum nans_sim_button_index {
    NANS_SIM_MOVE_FORWARD = 0,
    NANS_SIM_MOVE_BACK,
    NANS_SIM_MOVE_LEFT,
    NANS_SIM_MOVE_RIGHT,
    NANS_SIM_SHOOT,
    NANS_SIM_DEBUG_UP,
    NANS_SIM_DEBUG_DOWN,
    NANS_SIM_DEBUG_LEFT,
    NANS_SIM_DEBUG_RIGHT,
    NANS_SIM_DEBUG_FORWARD,
    NANS_SIM_DEBUG_BACK,
    NANS_SIM_DEBUG_RESET,
    NANS_SIM_DEBUG_CONTINUE,
    NANS_SIM_BUTTON_COUNT
};

typedef struct nans_sim_input {
    nans_sim_button Buttons[NANS_SIM_BUTTON_COUNT];
    float Sensitivity;
    int32_t XRel, YRel;
    int32_t X, Y;
} nans_sim_input;

/* The part of sdl_camera (code/nans.h:279-293) that input and shooting read. */
typedef struct nans_sim_camera {
    float position[3];
    float front[3];
    float up[3];
    float yaw, pitch, speed;
} nans_sim_camera;

#define NANS_SIM_MAGIC 0x4e53494du  /* "NSIM" */
#define NANS_SIM_MAX_BODIES 16384   /* the plain path's per-world limit */

/* Layout of PermanentStorage.  Vector arrays are [3][n] planes with
 * n = nboxes + nspheres (boxes first, as Cubes[] then Spheres[]); they hold the
 * state after the last frame.  The first frame's world is the reference's Init
 * scene (code/nans.cpp:1597-1678), or the single-world scene file named by the
 * environment variable NANS_SIM_SCENE. */
typedef struct nans_sim_state {
    uint32_t magic;
    int32_t status;               /* NANS_OK, or the code of the call that failed;
                                     a failed state stays failed */
    char error[256];
    nans_ctx *gpu;                /* device context, kept across reloads */
    int32_t uploads;              /* uploads of the world (1 after the first frame) */
    int32_t frames;               /* frames stepped */
    int32_t nboxes, nspheres;
    nans_params params;           /* params.dt follows the frame's dt */
    nans_sim_camera camera;
    float pos[3 * NANS_SIM_MAX_BODIES];
    float vel[3 * NANS_SIM_MAX_BODIES];
    float angvel[3 * NANS_SIM_MAX_BODIES];
    float angles[3 * NANS_SIM_MAX_BODIES];
    float force[3 * NANS_SIM_MAX_BODIES];
    float torque[3 * NANS_SIM_MAX_BODIES];
    float scale[3 * NANS_SIM_MAX_BODIES];
    float mass[NANS_SIM_MAX_BODIES];
    float moi[NANS_SIM_MAX_BODIES];
    int32_t kind[NANS_SIM_MAX_BODIES];
} nans_sim_state;

/* SIM_UPDATE_AND_RENDER (code/nans.h:396-397), bound by the platform with
 * dlsym(lib, "SimUpdateAndRender") (code/sdl_nans.cpp:404-405).  A failure is
 * printed once and recorded in state->status / state->error; later frames of a
 * failed state do nothing. */
void SimUpdateAndRender(nans_sim_memory *Memory, nans_sim_input *Input, void *Render, float dt);

/* Destroys the device context kept in PermanentStorage and clears IsInitialized.
 * No counterpart in the reference, whose platform exits without a shutdown call. */
void nans_sim_release(nans_sim_memory *Memory);

#ifdef __cplusplus
}
#endif

#endif
