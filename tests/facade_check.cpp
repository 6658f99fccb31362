// facade_check.cpp -- drives include/stt.hpp (the C++ facade with the
// reference's class names) for tests/test_facade.py.
//   facade_check errors            failure paths (no device needed)
//   facade_check producer DIR ...  one view's forward projection and ray-BP
//                                  scatter through TypicalRayProducer
//   facade_check render DIR ...    ST_Tomography::SetData + Render
// Test infrastructure: prints one line per check.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "stt.hpp"

namespace {

int failures = 0;

void expect_error(const char* what, int code, void (*fn)()) {
    try {
        fn();
        std::printf("FAIL %s: no error\n", what);
        ++failures;
    } catch (const stt::Error& e) {
        if (e.code == code) {
            std::printf("OK %s: %d %s\n", what, e.code, e.what());
        } else {
            std::printf("FAIL %s: code %d (%s), expected %d\n", what, e.code, e.what(), code);
            ++failures;
        }
    }
}

std::vector<float> read_raw(const std::string& path, size_t n) {
    std::vector<float> v(n);
    std::ifstream f(path, std::ios::binary);
    f.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(n * sizeof(float)));
    if (!f) throw stt::Error(STT_ERR_INVALID, "cannot read " + path);
    return v;
}

void write_raw(const std::string& path, const float* p, size_t n) {
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char*>(p), (std::streamsize)(n * sizeof(float)));
}

int errors(bool have_device) {
    expect_error("VolumeData with no frames", STT_ERR_INVALID, [] {
        const int d[3] = {8, 8, 8};
        const float s[3] = {1, 1, 1};
        VolumeData v(d, s, 0);
    });
    {
        const int d[3] = {6, 5, 4};
        const float s[3] = {1, 1, 1};
        VolumeData v(d, s, 3);
        const bool ok = v.m_Volumelist.size() == 4 && v.m_Flowlist.size() == 3 && v.m_Volumelist[3].size() == 120 &&
                        v.m_Flowlist[2].size() == 360;
        std::printf("%s VolumeData holds F+1 volumes and F flows\n", ok ? "OK" : "FAIL");
        failures += !ok;
    }
    {
        TypicalRayProducer rp(8, 9, 10, 16, 12, 1.0, "Voxelbased");
        const int d[3] = {12, 11, 10};
        const float s[3] = {0.5f, 1.0f, 2.0f};
        rp.SetVolSize(d, s);
        int d2[3];
        float s2[3];
        rp.GetVolSize(d2, s2);
        int img[2];
        rp.GetImageInUseSize(img);
        const bool ok = d2[0] == 12 && d2[2] == 10 && s2[2] == 2.0f && img[0] == 16 && img[1] == 12;
        std::printf("%s TypicalRayProducer volume and image sizes\n", ok ? "OK" : "FAIL");
        failures += !ok;
    }
    expect_error("RenderImage before UpdateMatrixAndBounds", STT_ERR_STATE, [] {
        TypicalRayProducer rp(8, 8, 8, 16, 16, 1.0, "Voxelbased");
        rp.RenderImage();
    });
    if (!have_device)
        expect_error("UpdateMatrixAndBounds without a device", STT_ERR_NODEVICE, [] {
            TypicalRayProducer rp(8, 8, 8, 16, 16, 1.0, "Voxelbased");
            rp.UpdateMatrixAndBounds();
        });
    // ST_Tomography reads lds.txt in its constructor, as the reference does
    std::remove("lds.txt");
    expect_error("ST_Tomography without lds.txt", STT_ERR_INVALID, [] {
        int wh[2] = {16, 16}, vol[3] = {8, 8, 8};
        double o[3] = {0, 0, 0}, sdg[1] = {0};
        const char* proj[1] = {"proj.tif"};
        ST_Tomography st(0.2, 0.2, 0.1, 0, 0, 0.1, "out/rec", wh, 10, 1, 1, 0.1, 600, proj, o, 1000, 1, "ATV",
                         "Voxelbased", vol, 2, sdg, 1);
    });
    {
        std::ofstream("lds.txt") << "0 7\n";
    }
    expect_error("SART before SetData", STT_ERR_STATE, [] {
        int wh[2] = {16, 16}, vol[3] = {8, 8, 8};
        double o[3] = {0, 0, 0}, sdg[1] = {0};
        const char* proj[1] = {"proj.tif"};
        ST_Tomography st(0.2, 0.2, 0.1, 0, 0, 0.1, "out/rec", wh, 10, 1, 1, 0.1, 600, proj, o, 1000, 1, "ATV",
                         "Voxelbased", vol, 2, sdg, 1);
        st.SART(0);
    });
    expect_error("SetData with the wrong frame count", STT_ERR_INVALID, [] {
        int wh[2] = {16, 16}, vol[3] = {8, 8, 8};
        double o[3] = {0, 0, 0}, sdg[1] = {0};
        const char* proj[1] = {"proj.tif"};
        ST_Tomography st(0.2, 0.2, 0.1, 0, 0, 0.1, "out/rec", wh, 10, 1, 1, 0.1, 600, proj, o, 1000, 1, "ATV",
                         "Voxelbased", vol, 2, sdg, 1);
        const int d[3] = {8, 8, 8};
        const float s[3] = {1, 1, 1};
        VolumeData v(d, s, 3);
        st.SetData(&v);
    });
    expect_error("SetData with another sample distance", STT_ERR_INVALID, [] {
        int wh[2] = {16, 16}, vol[3] = {8, 8, 8};
        double o[3] = {0, 0, 0}, sdg[1] = {0};
        const char* proj[1] = {"proj.tif"};
        ST_Tomography st(0.2, 0.2, 0.1, 0, 0, 0.1, "out/rec", wh, 10, 1, 1, 0.1, 600, proj, o, 1000, 1, "ATV",
                         "Voxelbased", vol, 2, sdg, 1);
        st.SetSampleDistance(1.0f);
        const int d[3] = {8, 8, 8};
        const float s[3] = {1, 1, 1};
        VolumeData v(d, s, 2);
        st.SetData(&v);
    });
    // a projection stack of 16x16 pages, 19 of them: 10 + 9 views, not 10 per frame
    {
        std::vector<float> stack(16 * 16 * 19, 1.0f);
        stt::check(stt_save_tiff("proj.tif", stack.data(), 16, 16, 19, 1), "save");
    }
    expect_error("SetData with a stack of the wrong size", STT_ERR_INVALID, [] {
        int wh[2] = {16, 16}, vol[3] = {8, 8, 8};
        double o[3] = {0, 0, 0}, sdg[1] = {0};
        const char* proj[1] = {"proj.tif"};
        ST_Tomography st(0.2, 0.2, 0.1, 0, 0, 0.1, "out/rec", wh, 10, 1, 1, 0.1, 600, proj, o, 1000, 1, "ATV",
                         "Voxelbased", vol, 2, sdg, 1);
        const int d[3] = {8, 8, 8};
        const float s[3] = {1, 1, 1};
        VolumeData v(d, s, 2);
        st.SetData(&v);
    });
    return failures;
}

// producer DIR nx ny nz W H ds sid sdd nview
int producer(char** a) {
    const std::string dir = a[0];
    const int nx = std::atoi(a[1]), ny = std::atoi(a[2]), nz = std::atoi(a[3]), W = std::atoi(a[4]),
              H = std::atoi(a[5]);
    const double ds = std::atof(a[6]), sid = std::atof(a[7]), sdd = std::atof(a[8]);
    const int nview = std::atoi(a[9]);
    const size_t V = (size_t)nx * ny * nz, D = (size_t)W * H;
    stt_geometry g{};
    g.vol[0] = nx;
    g.vol[1] = ny;
    g.vol[2] = nz;
    g.spacing[0] = g.spacing[1] = g.spacing[2] = 1.0f;
    g.det[0] = W;
    g.det[1] = H;
    g.ds = ds;
    g.sid = sid;
    g.sdd = sdd;
    const std::vector<float> vol = read_raw(dir + "/vol.raw", V), q = read_raw(dir + "/q.raw", 4 * (size_t)nview),
                             corr = read_raw(dir + "/corr.raw", D);
    TypicalRayProducer fp(nx, ny, nz, W, H, ds, "Voxelbased");
    TypicalRayProducer bp(nx, ny, nz, W, H, ds, "Raybased");
    fp.SetVolumeData(vol.data());
    bp.SetVolumeData(vol.data());
    for (int k = 0; k < nview; ++k) {
        float view[16], proj[16];
        stt::check(stt_scene_matrices(&g, q.data() + 4 * k, view, proj), "scene");
        fp.SetViewMatrix(view);
        fp.SetProjectionMatrix(proj);
        fp.UpdateMatrixAndBounds();
        fp.RenderImage();
        const std::string n = std::to_string(k);
        write_raw(dir + "/ax" + n + ".raw", fp.GetImageData(), D);
        write_raw(dir + "/rl" + n + ".raw", fp.GetImageRaylength(), D);
        write_raw(dir + "/w" + n + ".raw", fp.GetWeightOfImage(), D);
        bp.Initial();
        bp.SetViewMatrix(view);
        bp.SetProjectionMatrix(proj);
        bp.UpdateMatrixAndBounds();
        std::copy(corr.begin(), corr.end(), bp.GetCorrection());
        bp.SetBackProjectionMode();
        bp.RenderImage();
        write_raw(dir + "/c" + n + ".raw", bp.GetContributions(), V);
        write_raw(dir + "/s" + n + ".raw", bp.GetTotalWeight(), V);
    }
    std::printf("OK producer %d views\n", nview);
    return 0;
}

// render DIR nx ny nz W H sid sdd k e outer passes
int render(char** a) {
    const std::string dir = a[0];
    int vol[3] = {std::atoi(a[1]), std::atoi(a[2]), std::atoi(a[3])};
    int wh[2] = {std::atoi(a[4]), std::atoi(a[5])};
    const double sid = std::atof(a[6]), sdd = std::atof(a[7]), k = std::atof(a[8]), e = std::atof(a[9]);
    const int outer = std::atoi(a[10]), passes = std::atoi(a[11]);
    const size_t V = (size_t)vol[0] * vol[1] * vol[2];
    double o[3] = {0, 0, 0}, sdg[1] = {0};
    const std::string projfile = dir + "/proj.tif";
    const char* proj[1] = {projfile.c_str()};
    ST_Tomography st(0.2, 0.5, 0.1, k, e, 0.1, (dir + "/out/rec").c_str(), wh, 10, 1, 1, 0.1, sid, proj, o, sdd, 1.0,
                     "ATV", "Voxelbased", vol, 2, sdg, 1);
    const float s[3] = {1, 1, 1};
    VolumeData data(vol, s, 2);
    for (int f = 0; f < 2; ++f) data.m_Volumelist[f] = read_raw(dir + "/init" + std::to_string(f) + ".raw", V);
    data.m_Volumelist[2] = data.m_Volumelist[1];
    st.SetData(&data);
    st.SetIterations(outer, passes);
    st.SetFlowDirectory((dir + "/FlowOutput").c_str());
    st.Render();
    std::printf("OK render\n");
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        const std::string mode = argc > 1 ? argv[1] : "";
        if (mode == "errors") return errors(argc > 2 && std::string(argv[2]) == "device") ? 1 : 0;
        if (mode == "producer" && argc == 12) return producer(argv + 2);
        if (mode == "render" && argc == 14) return render(argv + 2);
        std::printf("usage: facade_check errors [device] | producer DIR ... | render DIR ...\n");
        return 2;
    } catch (const std::exception& e) {
        std::printf("FAIL exception: %s\n", e.what());
        return 1;
    }
}
