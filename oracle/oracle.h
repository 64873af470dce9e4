// ORACLE — test infrastructure only. C API of the CPU restatement, loaded by tests/
// and bench.py (cpu_baseline) through ctypes. The product never links it.
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

// Caller-allocated export arrays of a canonical LBVH (n primitives).
typedef struct oracle_bvh_out {
    uint32_t* codes_sorted;  // n
    int32_t* perm;           // n   sorted position -> primitive number
    int32_t* left;           // n-1 unified child index
    int32_t* right;          // n-1
    float* int_bounds;       // 6(n-1) min xyz, max xyz
    int32_t* leaf_prim;      // n   primitive referenced by leaf k
    float* leaf_bounds;      // 6n
    int32_t* parent;         // 2n-1 parent of every node in the unified index space (-1 at the root)
    float scene_bounds[6];
} oracle_bvh_out;

uint32_t oracle_morton3(float x, float y, float z);
int oracle_lbvh_build(const float* tri_xyz, int32_t n, int32_t mode, oracle_bvh_out* out);

// Philox / sampler access for known-answer tests.
void oracle_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
float oracle_sample(uint32_t seed, uint32_t pixel, uint32_t sample, int32_t dim);

// Render parameters, mirroring lts_render_params where meaningful.
typedef struct oracle_render_params {
    int32_t spp;
    int32_t max_bounce;
    int32_t rr_start;
    uint32_t seed;
    int32_t stratified;   // 1: film dimensions use Sampler::stratifiedSample
    int32_t threads;      // 0: hardware concurrency
    int32_t row_begin;    // render only rows [row_begin, row_end) of the film (bounded CPU samples)
    int32_t row_end;      // <= 0: all rows
} oracle_render_params;

typedef struct oracle_render_stats {
    double seconds;       // wall time of the render loop only
    int64_t samples;      // camera samples traced
    int64_t node_visits;  // BVH nodes whose box was tested
    int64_t tri_tests;    // ray-triangle tests
    int64_t rays;         // rays traced (closest-hit + any-hit + light re-intersections excluded)
    int32_t threads;
} oracle_render_stats;

typedef struct oracle_scene oracle_scene;

// Loads <project_dir>/<scene>.scene (+ <subject>.subject when non-empty), as the reference's
// parseScene does. Returns NULL on failure (message via oracle_last_error()).
oracle_scene* oracle_load_scene(const char* project_dir, const char* scene, const char* subject,
                                int32_t width, int32_t height);
void oracle_free_scene(oracle_scene* s);
const char* oracle_last_error(void);
int32_t oracle_scene_info(const oracle_scene* s, int32_t* n_tris, int32_t* n_lights, int32_t* n_meshes);

// Renders and writes the final image (getRender row order, W*H*3 floats) to rgb_out.
// If sample_rgb is non-NULL it also receives each sample's radiance L before filtering,
// laid out [sample][y][x][3].
int oracle_render(oracle_scene* s, const oracle_render_params* p, float* rgb_out,
                  float* sample_rgb, oracle_render_stats* st);

// Single-ray query for unit tests: closest hit (returns prim or -1, t) and any hit.
int32_t oracle_intersect(oracle_scene* s, const float o[3], const float d[3], float tmax, float* t_out);
int32_t oracle_occluded(oracle_scene* s, const float o[3], const float d[3], float tmax);

// Camera ray at film position (px, py) with a centred lens sample, differentials scaled for spp
// samples: out = o, d, rxO, rxD, ryO, ryD (18 floats).
void oracle_camera_ray(oracle_scene* s, float px, float py, int32_t spp, float* out);
// Primary hit of that ray and its differentials: out = p, n, uv, dpdx, dpdy, du/dx, dv/dx, du/dy,
// dv/dy (18 floats); returns the triangle or -1.
int32_t oracle_primary_hit(oracle_scene* s, float px, float py, int32_t spp, float* out);
// InfiniteLight::sampleLi from point p with sample u: out = wi, pdf. -1 without an infinite light.
int32_t oracle_infinite_sample(oracle_scene* s, const float p[3], const float u[2], float* out);
float oracle_infinite_pdf(const float wi[3]);
// Light l: out = kind, power entry, centre (3), radius, area, triangle (8 floats).
int32_t oracle_light_info(oracle_scene* s, int32_t l, float* out);
// Light pick by Distribution1D::sampleDiscrete over the light powers.
int32_t oracle_sample_light(oracle_scene* s, float u, float* pdf);

// Gaussian filter table (Filter.cuh:32-50, 82-99): 16x16 floats.
void oracle_filter_table(float* out256);
// The deterministic film gather (Camera::addSample, Camera.cu:192-218, as a gather): pfilm holds
// spp x h x w film positions (x, y), L the radiance of each (3 floats); film receives h x w x
// {r, g, b, weight} sums.
void oracle_film_gather(int32_t w, int32_t h, int32_t spp, const float* pfilm, const float* L, float* film);
// Texture pipeline of scenes loaded afterwards: 0 the reference's mipmap, 1 the filtered one.
void oracle_set_texture_filter(int32_t filtered);
void oracle_mip_lookup(const float* rgb, int32_t w, int32_t h, int32_t wrap, int32_t filtered, float s, float t,
                       float width, float* out3);

#ifdef __cplusplus
}
#endif

#ifdef __cplusplus
extern "C" {
#endif
// Component entry points for unit tests of the restatement.
float oracle_perlin(float x, float y, float z);
float oracle_fbm(const float p[3], const float dpdx[3], const float dpdy[3], float omega, int32_t octaves);
float oracle_turbulence(const float p[3], const float dpdx[3], const float dpdy[3], float omega, int32_t octaves);
void oracle_marble(const float p[3], int32_t octaves, float omega, float scale, float variation, float rgb[3]);
float oracle_fr_dielectric(float cos_theta_i, float eta_i, float eta_t);
void oracle_transform(const float t[3], const float axis[3], float angle, float scale, const float p[3], const float n[3],
                      float* out);
// BxDF in its local frame: kind 0 Lambert, 1 Oren-Nayar (param = sigma in degrees),
// 3 FresnelSpecular (param = eta), 4 microfacet reflection (param = roughness); R = 1.
// mode 0: f(wo, wi)[0] -> out[0]; 1: pdf(wo, wi) -> out[0];
// 2: sample with u -> out = {wi.x, wi.y, wi.z, pdf, f.r, sampledType}
void oracle_bxdf(int32_t kind, float param, int32_t mode, const float wo[3], const float wi[3], const float u[2], float* out);
float oracle_roughness_to_alpha(float roughness);

// Learned radiance cache network (nrc.cpp). Parameters: 5 hidden 64x64 matrices W[in][out]
// followed by the 64x16 output matrix (columns 3..15 padding, held at 0), 21504 floats.
// Raw record (16 floats): normalised position (3), outgoing direction theta/pi, phi/2pi (2),
// normal theta/pi, phi/2pi (2), roughness (1), diffuse reflectance (3), specular reflectance (3), 2 unused.
int32_t oracle_nrc_param_count(void);
void oracle_nrc_encode(const float raw[16], float out[64]);
// rounding: 0 none (fp64), 1 bf16, 2 fp16 at the GPU kernels' rounding points.
void oracle_nrc_forward(const float* params, const float* raw, int32_t n, int32_t rounding, double* y);
double oracle_nrc_grad(const float* params, const float* raw, const float* target, int32_t n, int32_t rounding,
                       const double* denom_ref, double* grad, double* y_out);
#ifdef __cplusplus
}
#endif
