// Host-side scene ingest: reads the reference's input files and produces the flat arrays of
// lts_scene.h, ready to upload.
#pragma once
#include <functional>
#include <string>
#include <vector>

#include "lts_scene.h"
#include "mesh_xf.h"

namespace lts {

struct HostScene {
    int width = 0, height = 0;
    CameraDesc cam{};
    float filter[256];
    int nMeshes = 0;
    std::vector<float> soup;        // 9 floats per primitive, global triangle order
    std::vector<ShadeTri> shade;    // per primitive
    std::vector<MatDesc> mats;
    std::vector<TexDesc> texs;
    std::vector<MipDesc> mips;
    std::vector<float> mipData;
    std::vector<LightDesc> lights;
    std::vector<float> lfunc, lcdf;
    float funcInt = 0;
    int infiniteLight = -1;
};

// A mesh's world transform applied to its points and normals (3 floats each) in place; false with
// `err` set on failure. The loader's transform runs on the GPU (mesh_transform).
using MeshTransformFn = std::function<bool(const MeshXf&, std::vector<float>& points, std::vector<float>& normals,
                                           std::string& err)>;

// parseScene (core/SceneLoader.cuh:195-360). Returns false with `err` set on failure.
// filteredTextures: build the filtered mip pyramids instead of the reference's.
bool load_scene_files(const MeshTransformFn& transform, const std::string& dir, const std::string& scene,
                      const std::string& subject, int width, int height, HostScene& out, std::string& err,
                      bool filteredTextures = false);

// OBJ import (core/FileIO.cuh:68-158, parseMeshFile with binMeshWrite): parses <stem>.obj and
// returns the bytes of the equivalent <stem>.bin (header nTri, nP, nUV, nN; points, uvs, normals;
// 9 indices per triangle). Returns false with `err` set on failure; *nTri = triangles read.
bool obj_mesh_bytes(const std::string& stem, std::vector<char>& bin, int* nTri, std::string& err);

// Gaussian filter table, radius (1, 1), alpha 1 (rendering/Filter.cuh:32-50, :82-99;
// LightTransportSimulation.cu:64).
void gaussian_filter_table(float* out256);

}  // namespace lts
