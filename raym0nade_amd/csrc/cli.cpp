// raym0nade_cli -- the reference's console (main.cpp:5-25, src/myconsole.cpp)
// on top of the C ABI.  Same command grammar and prompts:
//   create model <id>   Wavefront OBJ + MTL, PNG albedo maps, Radiance HDR sky
//   create scene <id> <config> <seed>   procedural workload 0/1/2/4
//   create args <id>    prompts exactly like MyConsole::createRenderArgs
//   delete model|args <id>, view model|args <id>, render <model> <args>, exit
// `render` is render_multiThread (src/render.cpp:593-677) with the GPU doing
// the pixel loop and the post path; the exports are written as
// savePath + "(" + tag + ").png" with Photo::save's 8-bit truncation.
#include <zlib.h>
#include <chrono>
#include <cstdio>
#include <iostream>
#include <limits>
#include <map>
#include <sstream>
#include <string>
#include <vector>
#include "../../include/raym0nade_hip.h"

namespace {

struct ModelEntry {
    rm_generated *gen = nullptr;
    rm_scene *scene = nullptr;
    std::string path;
    int64_t faces = 0;
};

struct ArgsEntry {
    rm_render_args a{};
    std::string savePath;
};

void put_be32(std::vector<unsigned char> &v, uint32_t x) {
    for (int s = 24; s >= 0; s -= 8) v.push_back((unsigned char)(x >> s));
}

void chunk(FILE *f, const char *type, const std::vector<unsigned char> &data) {
    std::vector<unsigned char> b;
    put_be32(b, (uint32_t)data.size());
    b.insert(b.end(), type, type + 4);
    b.insert(b.end(), data.begin(), data.end());
    uLong crc = crc32(0L, b.data() + 4, (uInt)(b.size() - 4));
    put_be32(b, (uint32_t)crc);
    fwrite(b.data(), 1, b.size(), f);
}

// Photo::save (src/image.cpp:481-532): 8-bit RGB, value = (unsigned char)(v * 255)
bool save_png(const std::string &path, const std::vector<float> &px, int w, int h) {
    std::vector<unsigned char> raw;
    raw.reserve((size_t)h * (w * 3 + 1));
    for (int y = 0; y < h; ++y) {
        raw.push_back(0);
        for (int x = 0; x < w; ++x)
            for (int k = 0; k < 3; ++k) raw.push_back(static_cast<unsigned char>(px[((size_t)y * w + x) * 3 + k] * 255));
    }
    uLongf clen = compressBound(raw.size());
    std::vector<unsigned char> comp(clen);
    if (compress2(comp.data(), &clen, raw.data(), raw.size(), 6) != Z_OK) return false;
    comp.resize(clen);
    FILE *f = fopen(path.c_str(), "wb");
    if (!f) {
        std::cerr << "Could not open file for writing" << std::endl;
        return false;
    }
    const unsigned char sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    fwrite(sig, 1, 8, f);
    std::vector<unsigned char> ihdr;
    put_be32(ihdr, (uint32_t)w);
    put_be32(ihdr, (uint32_t)h);
    ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});
    chunk(f, "IHDR", ihdr);
    chunk(f, "IDAT", comp);
    chunk(f, "IEND", {});
    fclose(f);
    return true;
}

class Console {
    std::map<std::string, ModelEntry> models;
    std::map<std::string, ArgsEntry> renderArgs;

public:
    ~Console() {
        for (auto &m : models) drop(m.second);
    }
    static void drop(ModelEntry &m) {
        if (m.scene) rm_scene_destroy(m.scene);
        if (m.gen) rm_generated_destroy(m.gen);
        m.scene = nullptr;
        m.gen = nullptr;
    }

    void createModel(const std::string &id) {
        if (models.count(id)) {
            std::cout << "Model (" << id << ") is already exists." << std::endl;
            return;
        }
        std::string folder, name, sky;
        std::cout << "Enter the model path (e.g., fbx/): ";
        std::cin >> folder;
        std::cout << "Enter the model name (e.g., model.fbx): ";
        std::cin >> name;
        std::cout << "Enter the sky map name (e.g., sky.hdr): ";
        std::cin >> sky;
        std::cin.ignore(std::numeric_limits<std::streamsize>::max(), '\n');
        ModelEntry m;
        if (rm_scene_load(folder.c_str(), name.c_str(), sky.c_str(), &m.scene)) {
            std::cout << "Error loading model: " << rm_last_error() << std::endl;
            return;
        }
        rm_scene_info info;
        rm_scene_get_info(m.scene, &info);
        std::cout << "faces: " << info.n_tris << std::endl;
        std::cout << "BVH has built with size " << info.n_nodes << std::endl;
        m.path = folder + name;
        m.faces = info.n_tris;
        models[id] = m;
    }

    void createScene(const std::string &id, int config, unsigned seed) {
        if (models.count(id)) {
            std::cout << "Model (" << id << ") is already exists." << std::endl;
            return;
        }
        ModelEntry m;
        if (rm_generate_scene(config, seed, nullptr, &m.gen)) {
            std::cout << "Error: " << rm_last_error() << std::endl;
            return;
        }
        rm_scene_desc d;
        rm_render_args a;
        rm_generated_desc(m.gen, &d, &a);
        if (rm_scene_create(&d, &m.scene)) {
            std::cout << "Error: " << rm_last_error() << std::endl;
            drop(m);
            return;
        }
        rm_scene_info info;
        rm_scene_get_info(m.scene, &info);
        std::cout << "faces: " << info.n_tris << std::endl;
        std::cout << "BVH has built with size " << info.n_nodes << std::endl;
        m.path = "procedural config " + std::to_string(config) + " seed " + std::to_string(seed);
        m.faces = info.n_tris;
        models[id] = m;
        // the workload's own camera is available as args "<id>_camera"
        ArgsEntry e;
        e.a = a;
        e.savePath = "output/" + id;
        renderArgs[id + "_camera"] = e;
        std::cout << "Model (" << id << ") created (args " << id << "_camera holds its camera)." << std::endl;
    }

    // MyConsole::createRenderArgs (src/myconsole.cpp:27-67)
    void createRenderArgs(const std::string &id) {
        if (renderArgs.count(id)) {
            std::cout << "Args (" << id << ") is already exists." << std::endl;
            return;
        }
        ArgsEntry e;
        rm_render_args &a = e.a;
        auto read = [](float v[3]) { std::cin >> v[0] >> v[1] >> v[2]; };
        std::cout << "direction (x,y,z): ";
        read(a.direction);
        std::cout << "right (x,y,z): ";
        read(a.right);
        std::cout << "up (x,y,z): ";
        read(a.up);
        std::cout << "position (D,R,U): ";
        float D, R, U;
        std::cin >> D >> R >> U;
        for (int k = 0; k < 3; ++k) a.position[k] = D * a.direction[k] + R * a.right[k] + U * a.up[k];
        std::cout << "accuracy, focus, CoC, exposure: ";
        std::cin >> a.accuracy >> a.focus >> a.coc >> a.exposure;
        std::cout << "width, height: ";
        std::cin >> a.width >> a.height;
        std::cout << "spp, threads, P_Direct: : ";
        std::cin >> a.spp >> a.threads >> a.p_direct;
        std::cout << "savePath: ";
        std::cin >> e.savePath;
        a.seed = 1;
        a.max_depth = 16;
        a.rr_min_depth = 1;
        renderArgs[id] = e;
        std::cout << "RenderArgs (" << id << ") created." << std::endl;
        std::cin.ignore(std::numeric_limits<std::streamsize>::max(), '\n');
    }

    void deleteModel(const std::string &id) {
        auto it = models.find(id);
        if (it == models.end()) {
            std::cout << "Model (" << id << ") does not exists." << std::endl;
            return;
        }
        drop(it->second);
        models.erase(it);
        std::cout << "Model (" << id << ") deleted." << std::endl;
    }
    void deleteRenderArgs(const std::string &id) {
        if (!renderArgs.erase(id)) {
            std::cout << "Args (" << id << ") does not exists." << std::endl;
            return;
        }
        std::cout << "RenderArgs (" << id << ") deleted." << std::endl;
    }
    void viewModel(const std::string &id) {
        auto it = models.find(id);
        if (it == models.end()) {
            std::cout << "Model (" << id << ") does not exists." << std::endl;
            return;
        }
        std::cout << "Model Path: " << it->second.path << std::endl;
        std::cout << "Faces: " << it->second.faces << std::endl;
    }
    void viewRenderArgs(const std::string &id) {
        auto it = renderArgs.find(id);
        if (it == renderArgs.end()) {
            std::cout << "Args (" << id << ") does not exists." << std::endl;
            return;
        }
        const rm_render_args &a = it->second.a;
        std::cout << "direction : " << a.direction[0] << " " << a.direction[1] << " " << a.direction[2] << std::endl;
        std::cout << "right : " << a.right[0] << " " << a.right[1] << " " << a.right[2] << std::endl;
        std::cout << "up : " << a.up[0] << " " << a.up[1] << " " << a.up[2] << std::endl;
        std::cout << "position : " << a.position[0] << " " << a.position[1] << " " << a.position[2] << std::endl;
        std::cout << "accuracy: " << a.accuracy << std::endl;
        std::cout << "exposure: " << a.exposure << std::endl;
        std::cout << "width, height: " << a.width << " " << a.height << std::endl;
        std::cout << "spp: " << a.spp << std::endl;
        std::cout << "threads: " << a.threads << std::endl;
        std::cout << "P_Direct: " << a.p_direct << std::endl;
        std::cout << "savePath: " << it->second.savePath << std::endl;
    }

    // render_multiThread (src/render.cpp:593-677)
    void render(const std::string &ms, const std::string &as) {
        auto mi = models.find(ms);
        if (mi == models.end()) {
            std::cout << "Model (" << ms << ") does not exists." << std::endl;
            return;
        }
        auto ai = renderArgs.find(as);
        if (ai == renderArgs.end()) {
            std::cout << "Args (" << as << ") does not exists." << std::endl;
            return;
        }
        const rm_render_args &a = ai->second.a;
        std::cout << "Rendering started on the GPU." << std::endl;
        auto t0 = std::chrono::steady_clock::now();
        rm_frame *frame = nullptr;
        if (rm_frame_create(a.width, a.height, &frame) || rm_render(mi->second.scene, &a, frame, nullptr)) {
            std::cout << "Error: " << rm_last_error() << std::endl;
            if (frame) rm_frame_destroy(frame);
            return;
        }
        std::vector<float> px((size_t)a.width * a.height * 3);
        rm_frame_download(frame, RM_BUF_PIXELS, px.data());  // synchronizes
        auto t1 = std::chrono::steady_clock::now();
        std::cout << "Rendering completed in " << std::chrono::duration<double, std::milli>(t1 - t0).count()
                  << " ms." << std::endl;
        const std::string &save = ai->second.savePath;
        auto exportImage = [&](const char *tag, int options) {
            if (rm_postprocess(frame, options, a.exposure, nullptr) || rm_frame_download(frame, RM_BUF_PIXELS, px.data())) {
                std::cout << "Error: " << rm_last_error() << std::endl;
                return;
            }
            save_png(save + "(" + tag + ").png", px, a.width, a.height);
        };
        // Photo::ShadeOption bits (include/image.h:17-34), in the order of
        // src/render.cpp:641-673 (bloom 256, FXAA 512, depth of field 1024)
        exportImage("DiffuseColor", 1);
        exportImage("DiffuseColor_FXAA", 1 | 512);
        exportImage("shapeNormal", 64);
        exportImage("surfaceNormal", 128);
        if (rm_spatial_clamp(frame, nullptr)) std::cout << "Error: " << rm_last_error() << std::endl;
        exportImage("Direct_Diffuse", 4 | 16);
        exportImage("Direct_Specular", 4 | 32);
        exportImage("Indirect_Diffuse", 8 | 16);
        exportImage("Indirect_Specular", 8 | 32);
        exportImage("Raw", 63);
        exportImage("Raw_Bloom", 63 | 256);
        exportImage("Raw_FXAA", 63 | 512);
        exportImage("Raw_Bloom_FXAA", 63 | 256 | 512);
        if (rm_filter(frame, nullptr)) std::cout << "Error: " << rm_last_error() << std::endl;
        exportImage("Direct_Diffuse_Filter", 4 | 16);
        exportImage("Direct_Specular_Filter", 4 | 32);
        exportImage("Indirect_Diffuse_Filter", 8 | 16);
        exportImage("Indirect_Specular_Filter", 8 | 32);
        exportImage("Filter", 63);
        exportImage("Filter_Bloom", 63 | 256);
        exportImage("Filter_FXAA", 63 | 512);
        exportImage("Filter_Bloom_FXAA", 63 | 256 | 512);
        if (a.coc > 1e-4f) {  // the frame holds focus, CoC and the camera position from rm_render
            exportImage("BaseColor_DepthFieldBlur", 1 | 1024);
            exportImage("Filter_DepthFieldBlur", 63 | 1024);
            exportImage("Filter_DepthFieldBlur_Bloom", 63 | 1024 | 256);
            exportImage("Filter_DepthFieldBlur_FXAA", 63 | 1024 | 512);
            exportImage("Filter_DepthFieldBlur_Bloom_FXAA", 63 | 1024 | 256 | 512);
        }
        rm_frame_destroy(frame);
        auto t2 = std::chrono::steady_clock::now();
        std::cout << "Post processing finished. Total: " << std::chrono::duration<double, std::milli>(t2 - t0).count()
                  << " ms." << std::endl;
    }
};

// parseCommand (src/myconsole.cpp:128-160) + `create scene`
void parseCommand(Console &c, const std::string &opt) {
    std::istringstream iss(opt);
    std::string command, type, name, m, a;
    iss >> command;
    if (command == "create") {
        iss >> type >> name;
        if (type == "model") c.createModel(name);
        else if (type == "args") c.createRenderArgs(name);
        else if (type == "scene") {
            int cfg = 0;
            unsigned seed = 1;
            iss >> cfg >> seed;
            c.createScene(name, cfg, seed);
        }
    } else if (command == "delete") {
        iss >> type >> name;
        if (type == "model" || type == "scene") c.deleteModel(name);
        else if (type == "args") c.deleteRenderArgs(name);
    } else if (command == "view") {
        iss >> type >> name;
        if (type == "model" || type == "scene") c.viewModel(name);
        else if (type == "args") c.viewRenderArgs(name);
    } else if (command == "render") {
        iss >> m >> a;
        c.render(m, a);
    } else {
        std::cout << "Unknown command." << std::endl;
    }
}

}  // namespace

int main() {
    Console console;
    std::string opt;
    while (true) {
        std::cout << "> ";
        if (!std::getline(std::cin, opt)) break;
        if (opt == "exit") break;
        if (opt.empty()) continue;
        parseCommand(console, opt);
    }
    return 0;
}
