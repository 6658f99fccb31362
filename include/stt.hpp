/*
 * stt.hpp — C++ facade of the C-ABI with the reference's class names.
 *
 * The reference (gmzang/SpaceTimeTomography) drives its reconstruction
 * through in-process classes: TypicalRayProducer (ImageFormationModel/
 * TypicalRayProducer.h:8-44), RayCaster (RayCaster.h:16), VolumeData
 * (VolumeData.h:24-100) and ST_Tomography (ST_Tomo/ST_Tomography.h:27-35,73).
 * This header re-creates that call surface on top of include/stt.h so a
 * caller written against those classes keeps its calls: header only, links
 * against libstt.so, and uses nothing but the C-ABI.
 *
 * Differences a caller sees:
 *   - Matrices are taken as anything with a public float ele[16] (the
 *     reference's Matrix4x4, column-major) or as a float[16].
 *   - Volumes and flows are std::vector<float> in CImg's layout (x fastest;
 *     a flow is three planes u, v, w) instead of CImg images.
 *   - Errors throw stt::Error (code + stt_last_error() text) where the
 *     reference printed and continued; no exception crosses the C-ABI.
 *   - The volume a TypicalRayProducer samples is set with SetVolumeData()
 *     (the reference's SingleScalarVolumePrepare, SingleScalarVolume.core.h:
 *     195-223, passed it through file-scope statics).
 *   - SART sweeps views in table order (the reference shuffles them with a
 *     time seed, ST_Tomography.cpp:635-642).
 * Define STT_FACADE_NO_GLOBAL_NAMES to keep the classes inside namespace stt.
 */
#ifndef STT_HPP
#define STT_HPP

#include <sys/stat.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "stt.h"

namespace stt {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& what) : std::runtime_error(what), code(c) {}
};

inline void check(int rc, const char* where) {
    if (rc != STT_OK) throw Error(rc, std::string(where) + ": " + stt_last_error());
}

/* VolumeData (VolumeData.h:24-100): F frame volumes plus the padding copy of
 * the last one (VolumeData.cpp:126) and F flows (the last one zero,
 * VolumeData.cpp:110). */
class VolumeData {
public:
    VolumeData(const int dims[3], const float spacings[3], int nframes) {
        if (nframes < 1 || dims[0] < 1 || dims[1] < 1 || dims[2] < 1)
            throw Error(STT_ERR_INVALID, "VolumeData: need positive dimensions and at least one frame");
        for (int c = 0; c < 3; ++c) {
            m_Dims[c] = dims[c];
            m_Spacings[c] = spacings[c];
        }
        const size_t V = (size_t)dims[0] * dims[1] * dims[2];
        m_Volumelist.assign((size_t)nframes + 1, std::vector<float>(V, 0.0f));
        m_Flowlist.assign((size_t)nframes, std::vector<float>(3 * V, 0.0f));
    }
    void GetDimensions(int dims[3]) const {
        for (int c = 0; c < 3; ++c) dims[c] = m_Dims[c];
    }
    void GetSpacings(float s[3]) const {
        for (int c = 0; c < 3; ++c) s[c] = m_Spacings[c];
    }
    /* The reference switched its volume accessor between sampling and
     * scattering (VolumeData.h:55); the library needs no switch. */
    void SetProjectionMode(int mode) { m_ProjectionMode = mode; }
    int nframes() const { return (int)m_Flowlist.size(); }

    std::vector<std::vector<float>> m_Volumelist;
    std::vector<std::vector<float>> m_Flowlist;

private:
    int m_Dims[3];
    float m_Spacings[3];
    int m_ProjectionMode = 0;
};

/* RayCaster (RayCaster.h:16): the sample distance.  The path samples at
 * |spacing| (ST_Tomography.cpp:300), the only value the reference sets;
 * ST_Tomography::SetData refuses any other. */
class RayCaster {
public:
    void SetSampleDistance(float sd) { m_sd = sd; }
    float GetSampleDistance() const { return m_sd; }

private:
    float m_sd = 0.0f;
};

/* TypicalRayProducer (TypicalRayProducer.h:8-44): one view's forward
 * projection (GenerateImage_MT, TypicalRayProducer.core.h:123-225) or, in
 * back-projection mode, the ray-driven scatter of the correction image. */
class TypicalRayProducer {
public:
    TypicalRayProducer(int volX, int volY, int volZ, int imgW, int imgH, double ds, const char* bp) {
        std::memset(&m_g, 0, sizeof(m_g));
        m_g.vol[0] = volX;
        m_g.vol[1] = volY;
        m_g.vol[2] = volZ;
        for (int c = 0; c < 3; ++c) m_g.spacing[c] = 1.0f;
        m_g.det[0] = imgW;
        m_g.det[1] = imgH;
        m_g.ds = ds;
        m_g.sid = 1.0;  // the camera comes in through the matrices
        m_g.sdd = 1.0;
        m_raybased = bp && std::strcmp(bp, "Raybased") == 0;
        const size_t D = (size_t)imgW * imgH;
        m_ImageData.assign(D, 0.0f);
        m_WeightOfImage.assign(D, 0.0f);
        m_imageraylength.assign(D, 0.0f);
        m_Correction.assign(D, 0.0f);
        resize_volume_buffers();
        for (int k = 0; k < 16; ++k) m_view[k] = m_proj[k] = (k % 5 == 0) ? 1.0f : 0.0f;
    }
    virtual ~TypicalRayProducer() {
        if (m_ctx) stt_ctx_destroy(m_ctx);
    }
    TypicalRayProducer(const TypicalRayProducer&) = delete;
    TypicalRayProducer& operator=(const TypicalRayProducer&) = delete;

    /* TypicalRayProducer::Initial (TypicalRayProducer.cpp:78-96). */
    void Initial() {
        m_BackProjection = false;
        std::fill(m_imageraylength.begin(), m_imageraylength.end(), 0.0f);
        std::fill(m_ImageData.begin(), m_ImageData.end(), 0.0f);
        std::fill(m_WeightOfImage.begin(), m_WeightOfImage.end(), 0.0f);
        std::fill(m_Correction.begin(), m_Correction.end(), 0.0f);
        if (m_raybased) {
            std::fill(m_EachVoxel_TotalWeight.begin(), m_EachVoxel_TotalWeight.end(), 0.0f);
            std::fill(m_Voxel_Contributions.begin(), m_Voxel_Contributions.end(), 0.0f);
        }
    }
    void SetBackProjectionMode() { m_BackProjection = true; }
    bool IsBackProjectionMode() const { return m_BackProjection; }

    void SetVolSize(const int volDims[3], const float volSpacings[3]) {
        for (int c = 0; c < 3; ++c) {
            m_g.vol[c] = volDims[c];
            m_g.spacing[c] = volSpacings[c];
        }
        resize_volume_buffers();
        drop_context();
    }
    void GetVolSize(int volDims[3], float volSpacings[3]) const {
        for (int c = 0; c < 3; ++c) {
            volDims[c] = m_g.vol[c];
            volSpacings[c] = m_g.spacing[c];
        }
    }
    /* The solver's detector distance (Scene::m_sdd), needed only by callers
     * that mix this producer with voxel-driven backprojection. */
    void SetDetectorDistance(double sdd) {
        m_g.sdd = sdd;
        drop_context();
    }

    void SetViewMatrix(const float* m16) { std::memcpy(m_view, m16, sizeof(m_view)); }
    void SetProjectionMatrix(const float* m16) { std::memcpy(m_proj, m16, sizeof(m_proj)); }
    // any matrix type with a public float ele[16] (the reference's Matrix4x4)
    template <class M, class = decltype(std::declval<const M&>().ele)>
    void SetViewMatrix(const M& m) { SetViewMatrix(m.ele); }
    template <class M, class = decltype(std::declval<const M&>().ele)>
    void SetProjectionMatrix(const M& m) { SetProjectionMatrix(m.ele); }

    /* TypicalRayProducer::UpdateMatrixAndBounds / _calculateMatrix
     * (TypicalRayProducer.cpp:142-208): the view record of the current
     * matrices (the context is created on the first call). */
    void UpdateMatrixAndBounds() {
        ensure_context();
        check(stt_set_view_matrices(m_ctx, 0, m_view, m_proj, 1), "UpdateMatrixAndBounds");
        m_ready = true;
    }

    /* The volume the rays sample (x fastest, volX*volY*volZ floats). */
    void SetVolumeData(const float* vol) {
        const size_t V = (size_t)m_g.vol[0] * m_g.vol[1] * m_g.vol[2];
        m_volume.assign(vol, vol + V);
        if (m_ctx) check(stt_set_volume(m_ctx, 0, m_volume.data()), "SetVolumeData");
    }

    /* GenerateImage_MT for the current view: line integrals, ray lengths
     * (missed rays keep theirs, :205-207) and sample counts; in
     * back-projection mode the scatter of GetCorrection() adds w*corr and w
     * into GetContributions() / GetTotalWeight() (SingleScalarVolume.core.h:
     * 73-84, 128-143). */
    void RenderImage() {
        if (!m_ready) throw Error(STT_ERR_STATE, "RenderImage: call UpdateMatrixAndBounds() first");
        if (m_volume.empty()) throw Error(STT_ERR_STATE, "RenderImage: call SetVolumeData() first");
        if (!m_BackProjection) {
            check(stt_forward_project(m_ctx, 0, 0, m_ImageData.data(), m_imageraylength.data(), m_WeightOfImage.data()),
                  "RenderImage");
            return;
        }
        if (!m_raybased) throw Error(STT_ERR_STATE, "RenderImage: back-projection mode needs a \"Raybased\" producer");
        const size_t V = m_volume.size();
        std::vector<float> c(V), w(V);
        // the library's ray-driven backprojection also normalises into its
        // copy of the volume; the sampled volume is restored after it
        check(stt_backproject_ray(m_ctx, 0, 0, m_Correction.data(), c.data(), w.data()), "RenderImage");
        check(stt_set_volume(m_ctx, 0, m_volume.data()), "RenderImage");
        for (size_t k = 0; k < V; ++k) {
            m_Voxel_Contributions[k] += c[k];
            m_EachVoxel_TotalWeight[k] += w[k];
        }
    }

    void GetImageInUseSize(int s[2]) const {
        s[0] = m_g.det[0];
        s[1] = m_g.det[1];
    }
    void GetImageOrigin(int o[2]) const { o[0] = o[1] = 0; }
    float* GetImageData() { return m_ImageData.data(); }
    float* GetWeightOfImage() { return m_WeightOfImage.data(); }
    float* GetImageRaylength() { return m_imageraylength.data(); }
    float* GetCorrection() { return m_Correction.data(); }
    float* GetContributions() { return m_Voxel_Contributions.data(); }
    float* GetTotalWeight() { return m_EachVoxel_TotalWeight.data(); }

private:
    void resize_volume_buffers() {
        const size_t V = m_raybased ? (size_t)m_g.vol[0] * m_g.vol[1] * m_g.vol[2] : 0;
        m_Voxel_Contributions.assign(V, 0.0f);
        m_EachVoxel_TotalWeight.assign(V, 0.0f);
    }
    void drop_context() {
        if (m_ctx) stt_ctx_destroy(m_ctx);
        m_ctx = nullptr;
        m_ready = false;
    }
    void ensure_context() {
        if (m_ctx) return;
        stt_params p;
        stt_params_default(&p);
        p.bp = m_raybased ? STT_BP_RAY : STT_BP_VOXEL;
        p.g_max = 3.0e38f;  // the producer alone never clamps (its normalisation is the caller's)
        const int one = 1;
        check(stt_ctx_create(&m_g, &p, 1, &one, &m_ctx), "TypicalRayProducer");
        if (!m_volume.empty()) check(stt_set_volume(m_ctx, 0, m_volume.data()), "TypicalRayProducer");
    }

    stt_geometry m_g;
    stt_ctx* m_ctx = nullptr;
    bool m_raybased = false, m_BackProjection = false, m_ready = false;
    float m_view[16], m_proj[16];
    std::vector<float> m_volume, m_ImageData, m_WeightOfImage, m_imageraylength, m_Correction;
    std::vector<float> m_Voxel_Contributions, m_EachVoxel_TotalWeight;
};

/* ST_Tomography (ST_Tomography.h:27-35; ctor ST_Tomography.cpp:200-279):
 * the space-time reconstruction with the reference's parameter list.  Its
 * effective values are the reference's: d1/d2 (sigma, tau) are unused there
 * (CP_sigma = 0.1125, CP_tau = 1), only the ATV prior exists, the rose ROI
 * cut[] applies, the start angles come from lds.txt (startdg is unused,
 * :258-267), and Render runs 20 outer iterations of 2 PD passes
 * (:95, :477). */
class ST_Tomography {
public:
    ST_Tomography(double d1, double d2, double d3, double d4, double d5, double d6, const char* outputfile,
                  int* imgwh, int projNo, int iters, int sartiters, double alp, double sid, const char** projsfile,
                  double* oxyz, double sdd, double ds, const char* prior, const char* bp, int* volxyz, int nframes,
                  double* startdg, int nrounds)
        : m_out(outputfile ? outputfile : "Test/test"), m_proj(projsfile && projsfile[0] ? projsfile[0] : ""),
          m_nframes(nframes), m_nrounds(nrounds), m_projno(projNo) {
        (void)d1;
        (void)d2;
        (void)iters;
        (void)sartiters;
        (void)prior;
        (void)startdg;
        if (nframes < 1 || nrounds < 1 || projNo < 1)
            throw Error(STT_ERR_INVALID, "ST_Tomography: nframes, nrounds and projNo must be positive");
        std::memset(&m_g, 0, sizeof(m_g));
        for (int c = 0; c < 3; ++c) {
            m_g.vol[c] = volxyz[c];
            m_g.spacing[c] = 1.0f;
            m_g.oxyz[c] = oxyz ? oxyz[c] : 0.0;
        }
        m_g.det[0] = imgwh[0];
        m_g.det[1] = imgwh[1];
        m_g.ds = ds;
        m_g.sid = sid;
        m_g.sdd = sdd;
        stt_params_reference(&m_p);
        m_p.lambda_tv = d3;
        m_p.k_tp = d4;
        m_p.e_bc = d5;
        m_p.huber = d6;
        m_p.alpha = alp;
        m_p.bp = bp && std::strcmp(bp, "Raybased") == 0 ? STT_BP_RAY : STT_BP_VOXEL;
        stt_of_params_reference(&m_ofp);
        m_start.assign((size_t)nframes * nrounds, 0.0f);
        check(stt_read_start_angles("lds.txt", nframes, nrounds, m_start.data()), "ST_Tomography");
    }
    virtual ~ST_Tomography() {
        if (m_ctx) stt_ctx_destroy(m_ctx);
    }
    ST_Tomography(const ST_Tomography&) = delete;
    ST_Tomography& operator=(const ST_Tomography&) = delete;

    void SetSampleDistance(float sd) { m_caster.SetSampleDistance(sd); }
    float GetSampleDistance() const { return m_caster.GetSampleDistance(); }

    /* SetData (ST_Tomography.cpp:287-361): the volume grid, the projection
     * stack split on z into frames (:355), the views of every frame
     * (InitViews :887-934: rounds of 10 views), the volumes and flows. */
    void SetData(VolumeData* data) {
        if (!data || data->nframes() != m_nframes)
            throw Error(STT_ERR_INVALID, "SetData: the VolumeData must hold nframes frames");
        m_data = data;
        float sp[3];
        data->GetDimensions(m_g.vol);
        data->GetSpacings(sp);
        for (int c = 0; c < 3; ++c) m_g.spacing[c] = sp[c];
        const float sd = std::sqrt(sp[0] * sp[0] + sp[1] * sp[1] + sp[2] * sp[2]);
        if (m_caster.GetSampleDistance() == 0.0f) m_caster.SetSampleDistance(sd);
        if (m_caster.GetSampleDistance() != sd)
            throw Error(STT_ERR_INVALID, "SetData: the path samples at |spacing| (ST_Tomography.cpp:300)");
        stt_image_info info;
        check(stt_image_query(m_proj.c_str(), &info), "SetData");
        if (info.width != m_g.det[0] || info.height != m_g.det[1] || info.spectrum != 1)
            throw Error(STT_ERR_INVALID, "SetData: the projection images do not match imgwh");
        m_views.assign(m_nframes, 0);
        check(stt_split_sizes(info.depth, m_nframes, m_views.data()), "SetData");
        for (int f = 0; f < m_nframes; ++f)
            if (m_views[f] != 10 * m_nrounds || m_views[f] != m_projno)
                throw Error(STT_ERR_INVALID, "SetData: every frame needs projNo = 10 views per round (InitViews)");
        if (m_ctx) stt_ctx_destroy(m_ctx);
        m_ctx = nullptr;
        check(stt_ctx_create(&m_g, &m_p, m_nframes, m_views.data(), &m_ctx), "SetData");
        const size_t D = (size_t)info.width * info.height;
        std::vector<float> stack(D * info.depth);
        check(stt_load_image(m_proj.c_str(), stack.data(), stack.size()), "SetData");
        size_t z0 = 0;
        for (int f = 0; f < m_nframes; ++f) {
            check(stt_set_projections(m_ctx, f, stack.data() + z0 * D), "SetData");
            z0 += (size_t)m_views[f];
            std::vector<float> q((size_t)4 * m_views[f]);
            check(stt_view_rounds(m_start.data() + (size_t)f * m_nrounds, m_nrounds, 10, q.data()), "SetData");
            check(stt_set_view_quaternions(m_ctx, f, q.data(), m_views[f]), "SetData");
        }
        for (int f = 0; f <= m_nframes; ++f) check(stt_set_volume(m_ctx, f, data->m_Volumelist[f].data()), "SetData");
        for (int f = 0; f + 1 < m_nframes; ++f) check(stt_set_flow(m_ctx, f, data->m_Flowlist[f].data()), "SetData");
        m_sart.assign(m_nframes, 1);
    }

    /* SART(iframes) (ST_Tomography.cpp:616-808): one sweep over the frame's
     * views, x *= imROI, the volume written to "<out>-%03d-frame-<n>iter.tiff"
     * and back into the VolumeData. */
    void SART(int frame) {
        ready();
        check(stt_sart_sweep(m_ctx, frame, nullptr, m_views.at(frame)), "SART");
        dump_frame(frame);
    }

    /* Render (ST_Tomography.cpp:397-614): per outer iteration the flow of
     * every frame pair (:421-469, dumped with the second frame warped back by
     * it), then from the second on the PD passes over all frames (:475-608),
     * each frame dumped after its sweep. */
    void Render() {
        ready();
        make_dirs(m_flow_dir);
        const int dims[3] = {m_g.vol[0], m_g.vol[1], m_g.vol[2]};
        const size_t V = (size_t)dims[0] * dims[1] * dims[2];
        std::vector<int> its((size_t)m_ofp.levels * m_ofp.warps);
        std::vector<float> uvw(3 * V), warped(V);
        for (int outer = 0; outer < m_outer; ++outer) {
            for (int f = 0; f + 1 < m_nframes; ++f) {
                check(stt_optical_flow(m_ctx, f, &m_ofp, its.data()), "Render");
                check(stt_get_buffer(m_ctx, STT_BUF_FLOW, f, uvw.data()), "Render");
                check(stt_warp_frame(m_ctx, f, -1.0f, STT_INTERP_CUBIC_CLAMPED, warped.data()), "Render");
                char base[512];
                std::snprintf(base, sizeof(base), "%s/%03d.%03d.%03d.frame%02d.iter%02d", m_flow_dir.c_str(), dims[0],
                              dims[1], dims[2], f, outer);
                check(stt_save_tiff((std::string(base) + ".tiff").c_str(), warped.data(), dims[0], dims[1], dims[2], 1),
                      "Render");
                check(stt_save_analyze((std::string(base) + ".field.hdr").c_str(), uvw.data(), dims[0], dims[1], dims[2],
                                       3, nullptr),
                      "Render");
                m_data->m_Flowlist[f] = uvw;
            }
            if (outer == 0) continue;
            if (outer == 1) check(stt_pd_init(m_ctx), "Render");
            for (int pass = 0; pass < m_passes; ++pass) {
                check(stt_iterate(m_ctx, 1, 1, 1, nullptr), "Render");
                for (int f = 0; f < m_nframes; ++f) dump_frame(f);
            }
        }
    }

    /* Loop counts and the flow output directory (the reference's constants:
     * 20 outer iterations, 2 PD passes, "FlowOutput"). */
    void SetIterations(int outer, int pd_passes) {
        m_outer = outer;
        m_passes = pd_passes;
    }
    void SetFlowDirectory(const char* dir) { m_flow_dir = dir; }
    stt_ctx* context() { return m_ctx; }

private:
    static void make_dirs(const std::string& dir) {
        for (size_t i = 1; i <= dir.size(); ++i)
            if (i == dir.size() || dir[i] == '/') (void)::mkdir(dir.substr(0, i).c_str(), 0755);
    }
    void ready() const {
        if (!m_ctx || !m_data) throw Error(STT_ERR_STATE, "call SetData() first");
    }
    void dump_frame(int f) {
        const size_t slash = m_out.rfind('/');
        if (slash != std::string::npos) make_dirs(m_out.substr(0, slash));
        check(stt_get_volume(m_ctx, f, m_data->m_Volumelist.at(f).data()), "SART");
        char name[1024];
        std::snprintf(name, sizeof(name), "%s-%03d-frame-%diter.tiff", m_out.c_str(), f, m_sart.at(f)++);
        check(stt_save_tiff(name, m_data->m_Volumelist[f].data(), m_g.vol[0], m_g.vol[1], m_g.vol[2], 1), "SART");
    }

    stt_geometry m_g;
    stt_params m_p;
    stt_of_params m_ofp;
    std::string m_out, m_proj, m_flow_dir = "FlowOutput";
    int m_nframes, m_nrounds, m_projno;
    int m_outer = 20, m_passes = 2;
    std::vector<float> m_start;
    std::vector<int> m_views, m_sart;
    RayCaster m_caster;
    VolumeData* m_data = nullptr;
    stt_ctx* m_ctx = nullptr;
};

}  // namespace stt

#ifndef STT_FACADE_NO_GLOBAL_NAMES
using stt::RayCaster;
using stt::ST_Tomography;
using stt::TypicalRayProducer;
using stt::VolumeData;
#endif

#endif
