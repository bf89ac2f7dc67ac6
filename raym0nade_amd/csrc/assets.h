// Host asset path shared by the model readers (assets.cpp: OBJ + MTL and the
// image decoders; fbx.cpp: binary FBX).  Everything here runs on the host and
// touches no device.
#pragma once
#include <cstdint>
#include <string>
#include <vector>
#include "../../include/raym0nade_hip.h"

namespace rm {

// A model read from disk: the arrays an rm_scene_desc points into (three
// corners per triangle, triangles grouped by mesh in mesh order).
struct LoadedModel {
    std::vector<float> verts, uvs, normals;
    std::vector<int32_t> material, mesh;
    std::vector<rm_material> mats;
    std::vector<std::string> mat_names;
    std::vector<std::vector<float>> tex_data;
    std::vector<std::vector<uint8_t>> tex_bytes;
    std::vector<rm_texture> texs;
    std::vector<std::string> tex_paths;
    std::vector<float> sky;
    int sky_w = 0, sky_h = 0;
    bool all_normals = true;
    rm_scene_desc desc{};  // filled by fill_desc
};

// A material and the file names of its maps, relative to the model folder
// (Material::texture[aiTextureType_*], src/material.cpp:290-298).
struct MtlDef {
    std::string name;
    rm_material m{};
    std::string map_kd, map_norm, map_ke, map_ks;
};

// Material() defaults: white without a diffuse map, roughness 0.8, opaque.
rm_material default_material();
// Material::loadMaterialProperties (src/material.cpp:300-333): the glass an
// opacity below 0.99 makes, with colours by material name.
void glass_by_name(MtlDef &d);
// Loads every map of defs (each file once) into S and appends the materials
// to S.mats / S.mat_names.
int32_t load_material_maps(const std::string &folder, std::vector<MtlDef> &defs, LoadedModel &S);
// Points S.desc at S's arrays.
void fill_desc(LoadedModel &S);

int32_t parse_obj(const std::string &folder, const std::string &name, LoadedModel &S);
int32_t parse_fbx(const std::string &folder, const std::string &name, LoadedModel &S);

}  // namespace rm
