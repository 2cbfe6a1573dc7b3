// C ABI: error reporting and the host-only entry points (scene generation
// and packing). Device entry points are in api.hip.
#include <cstdlib>
#include <new>
#include <string>

#include "../../include/photon_tracer.h"
#include "host.hpp"
#include "status.hpp"

namespace pt {

thread_local std::string g_last_error;

pt_status fail(pt_status s, const std::string& msg) {
    g_last_error = msg;
    return s;
}

uint32_t* copy_out(const std::vector<uint32_t>& v, size_t* n) {
    uint32_t* p = static_cast<uint32_t*>(std::malloc(v.size() * 4 + 4));
    if (!p) throw std::bad_alloc();
    if (!v.empty()) std::memcpy(p, v.data(), v.size() * 4);
    *n = v.size();
    return p;
}

}  // namespace pt

extern "C" {

const char* pt_last_error(void) { return pt::g_last_error.c_str(); }

const char* pt_version(void) { return "photon_tracer_amd 0.1 (gfx950)"; }

void pt_free(void* p) { std::free(p); }

pt_status pt_generate(int config, uint64_t seed, double scale, uint32_t** ptsd, size_t* n_words) {
    if (!ptsd || !n_words) return pt::fail(PT_ERR_INVALID, "pt_generate: null output pointer");
    PT_GUARD({
        const pt::scene_desc s = pt::generate_config(config, seed, scale);
        *ptsd = pt::copy_out(pt::ptsd_write(s), n_words);
    });
}

pt_status pt_generate_job(uint64_t seed, double scale, pt_job_desc* desc, uint32_t** ptsd, size_t* n_words) {
    if (!desc || !ptsd || !n_words) return pt::fail(PT_ERR_INVALID, "pt_generate_job: null output pointer");
    PT_GUARD({
        pt::job_desc d{};
        const pt::scene_desc s = pt::generate_job(seed, scale, d);
        desc->kind = d.kind;
        desc->width = d.width;
        desc->height = d.height;
        desc->spp = d.spp;
        desc->primitives = d.primitives;
        desc->seed = d.seed;
        *ptsd = pt::copy_out(pt::ptsd_write(s), n_words);
    });
}

pt_status pt_ptsd_add_obj(const uint32_t* ptsd, size_t n_words, const char* obj_text, size_t text_len,
                          const char* mtl_text, size_t mtl_len, const float base[3], double height, uint32_t material,
                          double turn_deg, uint32_t** out, size_t* out_words) {
    if (!ptsd || !out || !out_words || !base || (text_len && !obj_text) || (mtl_len && !mtl_text))
        return pt::fail(PT_ERR_INVALID, "pt_ptsd_add_obj: null pointer");
    if (!(height > 0)) return pt::fail(PT_ERR_INVALID, "pt_ptsd_add_obj: height must be > 0");
    PT_GUARD({
        pt::scene_desc s = pt::ptsd_read(ptsd, n_words);
        const std::string mtl(mtl_text ? mtl_text : "", mtl_len);
        pt::add_obj(s, std::string(obj_text ? obj_text : "", text_len), pt::v3(base[0], base[1], base[2]), height,
                    material, turn_deg, mtl_text ? &mtl : nullptr);
        *out = pt::copy_out(pt::ptsd_write(s), out_words);
    });
}

pt_status pt_ptsd_add_obj_textured(const uint32_t* ptsd, size_t n_words, const char* obj_text, size_t text_len,
                                   const char* mtl_text, size_t mtl_len, const pt_image* images, size_t n_images,
                                   const float base[3], double height, uint32_t material, double turn_deg,
                                   uint32_t** out, size_t* out_words) {
    if (!ptsd || !out || !out_words || !base || (text_len && !obj_text) || (mtl_len && !mtl_text) ||
        (n_images && !images))
        return pt::fail(PT_ERR_INVALID, "pt_ptsd_add_obj_textured: null pointer");
    if (!(height > 0)) return pt::fail(PT_ERR_INVALID, "pt_ptsd_add_obj_textured: height must be > 0");
    PT_GUARD({
        std::vector<pt::image_rec> ims(n_images);
        for (size_t k = 0; k < n_images; ++k) {
            const pt_image& in = images[k];
            if (!in.name || !in.rgb || in.width < 1 || in.height < 1 || in.width > 65536 || in.height > 65536)
                throw pt::status_error(PT_ERR_INVALID, "pt_ptsd_add_obj_textured: image without a name, pixels "
                                                       "or a size of 1 to 65536 a side");
            ims[k].name = in.name;
            ims[k].width = in.width;
            ims[k].height = in.height;
            ims[k].rgb.assign(in.rgb, in.rgb + size_t(in.width) * in.height * 3);
        }
        pt::scene_desc s = pt::ptsd_read(ptsd, n_words);
        const std::string mtl(mtl_text ? mtl_text : "", mtl_len);
        pt::add_obj(s, std::string(obj_text ? obj_text : "", text_len), pt::v3(base[0], base[1], base[2]), height,
                    material, turn_deg, mtl_text ? &mtl : nullptr, &ims);
        *out = pt::copy_out(pt::ptsd_write(s), out_words);
    });
}

pt_status pt_decode_image(const void* bytes, size_t n_bytes, uint32_t* width, uint32_t* height, uint8_t** rgb) {
    if (!width || !height || !rgb || (n_bytes && !bytes)) return pt::fail(PT_ERR_INVALID, "pt_decode_image: null pointer");
    *rgb = nullptr;
    PT_GUARD({
        try {
            const pt::image_rec img = pt::decode_image(static_cast<const uint8_t*>(bytes), n_bytes);
            uint8_t* p = static_cast<uint8_t*>(std::malloc(img.rgb.size()));
            if (!p) throw std::bad_alloc();
            std::memcpy(p, img.rgb.data(), img.rgb.size());
            *width = img.width;
            *height = img.height;
            *rgb = p;
        } catch (const std::runtime_error& e) {
            throw pt::status_error(PT_ERR_INVALID, std::string("pt_decode_image: ") + e.what());
        }
    });
}

pt_status pt_pack(const uint32_t* ptsd, size_t n_words, uint32_t** ptgp, size_t* ptgp_words) {
    if (!ptsd || !ptgp || !ptgp_words) return pt::fail(PT_ERR_INVALID, "pt_pack: null pointer");
    PT_GUARD({
        const pt::scene_desc s = pt::ptsd_read(ptsd, n_words);
        *ptgp = pt::copy_out(pt::pack_ptgp(s), ptgp_words);
    });
}

}  // extern "C"
