// Exercises the header-only GCT:: drop-in API over the C ABI.
//   api_test            geometry known answers (no GPU needed)
//   api_test --gpu DIR  also reconstructs a flat-panel scan through
//                       GCT::Recon::reconstruct and writes DIR/vol.raw
// Known answers follow the reference's tests:
//   test/cpp/geom/proj_geom_factory_params_test.cpp (views at 270 and 0 deg),
//   test/cpp/geom/rotation_geom_test.cpp (rotation analysis),
//   test/cpp/recon/check_validity_test.cpp (isValidAxialRecon).
#include "gct/gct.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <unistd.h>
#include <numbers>
#include <string>

using namespace GCT;
using namespace GCT::Geom;
using namespace GCT::Geom::Units;

static int failures = 0;
#define CHECK(c)                                                             \
    do {                                                                     \
        if (!(c)) {                                                          \
            std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #c);          \
            ++failures;                                                      \
        }                                                                    \
    } while (0)

static bool near(float a, float b, float eps = 1e-4f) { return std::fabs(a - b) <= eps; }

static RotationParams rot4()
{
    RotationParams r;
    r.dir = RotDir::CW;
    r.init_ang = Radians{std::numbers::pi_v<float> / 2.0f};
    r.ang_incr = Radians{std::numbers::pi_v<float> / 2.0f};
    r.num_proj = 4u;
    return r;
}

int main(int argc, char **argv)
{
    // 1-D parallel: third view at 270 deg, fourth at 360 deg (exact float sums)
    {
        Parallel1DProjGeomParams p;
        p.rot = rot4();
        p.det_row_center = 256.0_pix;
        auto pg2 = ProjGeomFactory<DetParallel>::create(p, 2u);
        auto pg3 = ProjGeomFactory<DetParallel>::create(p, 3u);
        CHECK(pg2.rot_ang.val == 3.0f * std::numbers::pi_v<float> / 2.0f);
        CHECK(pg3.rot_ang.val == 2.0f * std::numbers::pi_v<float>);
        CHECK(pg2.det_row_center.val == 256.0f);
    }
    // fan arc: at 270 deg the source is on -y, rows along +x, iso-ray along +y
    {
        FanBeamProjGeomParams p;
        p.rot = rot4();
        p.det_row_center = 256.0_pix;
        p.src_iso_dist = 500.0_mm;
        p.src_det_dist = 1000.0_mm;
        auto pg = ProjGeomFactory<DetFanArc>::create(p, 2u);
        CHECK(near(pg.src_pos.x.val, 0.0f) && near(pg.src_pos.y.val, -500.0f));
        CHECK(near(pg.det_row_dir.x.val, 1.0f) && near(pg.det_row_dir.y.val, 0.0f));
        CHECK(near(pg.iso_ray_dir.x.val, 0.0f) && near(pg.iso_ray_dir.y.val, 1.0f));
        auto pg0 = ProjGeomFactory<DetFanArc>::create(p, 3u);
        CHECK(near(pg0.src_pos.x.val, 500.0f) && near(pg0.src_pos.y.val, 0.0f));
        CHECK(near(pg0.det_row_dir.x.val, 0.0f) && near(pg0.det_row_dir.y.val, 1.0f));
        CHECK(near(pg0.iso_ray_dir.x.val, -1.0f) && near(pg0.iso_ray_dir.y.val, 0.0f));
    }
    // flat cone: at 270 deg, rows along +x, columns along -z
    {
        ConeBeamProjGeomParams p;
        p.rot = rot4();
        p.det_row_center = 256.0_pix;
        p.det_col_center = 257.0_pix;
        p.src_iso_dist = 500.0_mm;
        p.src_det_dist = 1000.0_mm;
        auto pg = ProjGeomFactory<DetFanLine, DetFanLine>::create(p, 2u);
        CHECK(near(pg.src_pos.x.val, 0.0f) && near(pg.src_pos.y.val, -500.0f) && near(pg.src_pos.z.val, 0.0f));
        CHECK(near(pg.det_row_dir.x.val, 1.0f) && near(pg.det_row_dir.y.val, 0.0f) && near(pg.det_row_dir.z.val, 0.0f));
        CHECK(near(pg.det_col_dir.x.val, 0.0f) && near(pg.det_col_dir.y.val, 0.0f) && near(pg.det_col_dir.z.val, -1.0f));
        CHECK(pg.det_col_center.val == 257.0f);
    }
    // rotation analysis of every geometry: CW, increment 90 deg, range 360 deg, 4 per rotation
    {
        ConeBeamProjGeomParams p;
        p.rot = rot4();
        p.det_row_center = 256.0_pix;
        p.det_col_center = 257.0_pix;
        p.src_iso_dist = 500.0_mm;
        p.src_det_dist = 1000.0_mm;
        std::vector<ProjGeom<DetFanLine, DetFanLine>> v;
        for (unsigned i = 0; i < 4; ++i) v.push_back(ProjGeomFactory<DetFanLine, DetFanLine>::create(p, i));
        AxialCTGeom<DetFanLine, DetFanLine> cg{std::make_tuple(DetFanLine{{512u, 1.0_mm}}, DetFanLine{{512u, 1.0_mm}}), v};
        auto const &rg = cg.rotationGeom();
        CHECK(rg.dir == RotDir::CW);
        CHECK(near(rg.rot_ang_incr.val, std::numbers::pi_v<float> / 2.0f));
        CHECK(near(rg.rot_ang_range.val, 2.0f * std::numbers::pi_v<float>));
        CHECK(rg.num_proj_per_rot == 4u);
        // isValidAxialRecon: matching and mismatching stacks
        Data::ProjectionStack<float, 2u> ok{{512u, 512u}, 4u};
        Data::ProjectionStack<float, 2u> bad_cols{{511u, 512u}, 4u};
        Data::ProjectionStack<float, 2u> bad_views{{512u, 512u}, 3u};
        CHECK(Recon::isValidAxialRecon(cg, ok));
        CHECK(!Recon::isValidAxialRecon(cg, bad_cols));
        CHECK(!Recon::isValidAxialRecon(cg, bad_views));
    }
    // a stationary trajectory throws from the constructor (ct_geom.h:26-32)
    {
        std::vector<ProjGeom<DetParallel>> v(3);
        for (auto &x : v) x.rot_ang = Radians{0.3f};
        bool threw = false;
        try {
            AxialCTGeom<DetParallel> cg{std::make_tuple(DetParallel{{16u, 1.0_mm}}), v};
        } catch (std::runtime_error const &) {
            threw = true;
        }
        CHECK(threw);
    }
    // zero-size image throws std::invalid_argument (recon_img.h:42-55)
    {
        bool threw = false;
        try {
            Data::ReconImg<float, 2u> im{ReconImgGeom<2u>{{0u, 1.0_mm}, {4u, 1.0_mm}, {}}};
        } catch (std::invalid_argument const &) {
            threw = true;
        }
        CHECK(threw);
    }

    // spiral geometry (spiral_ct_geom.h): 3 rotations of 64 views on a helix of 6 mm per rotation
    {
        ConeBeamProjGeomParams p;
        p.rot.init_ang = Radians{0.0f};
        p.rot.ang_incr = Radians{2.0f * std::numbers::pi_v<float> / 64.0f};
        p.rot.num_proj = 192u;
        p.rot.hel_pitch = 6.0f / (2.0f * std::numbers::pi_v<float>);
        p.det_row_center = 47.5_pix;
        p.det_col_center = 7.5_pix;
        p.src_iso_dist = 200.0_mm;
        p.src_det_dist = 400.0_mm;
        std::vector<ProjGeom<DetFanArc, DetFanLine>> v;
        for (unsigned i = 0; i < 192u; ++i) v.push_back(ProjGeomFactory<DetFanArc, DetFanLine>::create(p, i));
        const auto dets = std::make_tuple(DetFanArc{{96u, Radians{0.0025f}}}, DetFanLine{{16u, 1.0_mm}});
        SpiralCTGeom<DetFanArc, DetFanLine> sg{dets, v};
        CHECK(sg.rotationGeom().num_proj_per_rot == 64u);
        const auto [f, l] = sg.firstAndLastProjIndices(v[100].src_pos.z.val);
        CHECK(f == 68u && l == 132u);
        // no pitch: the constructor throws (spiral pitch cannot be zero)
        p.rot.hel_pitch = 0.0f;
        std::vector<ProjGeom<DetFanArc, DetFanLine>> c;
        for (unsigned i = 0; i < 192u; ++i) c.push_back(ProjGeomFactory<DetFanArc, DetFanLine>::create(p, i));
        bool threw = false;
        try {
            SpiralCTGeom<DetFanArc, DetFanLine> bad{dets, c};
        } catch (std::runtime_error const &) {
            threw = true;
        }
        CHECK(threw);
    }
    // array files through GCT::IO (meta_img_file_writer_tests.cpp, mhd_file_reader_tests.cpp):
    // the 3 x 2 x 4 iota fixture written as .mha and .mhd and read back, a series, a DEN image
    {
        namespace fs = std::filesystem;
        const fs::path dir = fs::temp_directory_path() / ("gct_io_" + std::to_string(::getpid()));
        fs::create_directories(dir);
        std::vector<float> data(24), data2(24);
        for (int i = 0; i < 24; ++i) { data[i] = float(i); data2[i] = float(24 + i); }
        IO::ArrayFileWriter::MetaImgMetaData<3u> md;
        md.dim_size = {3u, 2u, 4u};
        for (const char *ext : {".mha", ".mhd"}) {
            const std::string f1 = (dir / (std::string("01") + ext)).string();
            const std::string f2 = (dir / (std::string("02") + ext)).string();
            IO::ArrayFileWriter::write(md, data, f1);
            IO::ArrayFileWriter::write(md, data2, f2);
            const IO::MetaImgFileReader r{f1};
            const auto m = r.readMetadata();
            CHECK(m.num_rows == 2u && m.num_cols == 3u && m.num_slices == 4u && m.type == IO::ArrayElemType::T_FLOAT);
            CHECK(r.readData<float>() == data);
            const auto all = IO::MetaImgFileReader{f1, true}.readData<float>();
            CHECK(all.size() == 48u && all[0] == 0.0f && all[47] == 47.0f);
            const auto proj = IO::ProjectionReader<IO::MetaImgFileReader>::read<float, 2u>(f1);
            CHECK(proj.projSize().num_cols == 3u && proj.projSize().num_rows == 2u && proj.numProj() == 4u);
        }
        CHECK(IO::ArrayFileWriter::Detail::MetaImage::typeStr<std::uint16_t>() == "MET_USHORT");
        CHECK(IO::ArrayFileWriter::Detail::MetaImage::numDataElems(std::array<unsigned, 3UL>{2u, 3u, 4u}) == 24UL);
        bool threw = false;
        try {
            IO::ArrayFileWriter::write(md, std::vector<float>(1), (dir / "invalid_dir" / "raw.mha").string());
        } catch (std::runtime_error const &) {
            threw = true;
        }
        CHECK(threw);
        Data::ReconImg<float, 2u> im{ReconImgGeom<2u>{{3u, 0.5_mm}, {2u, 0.5_mm}, {}}};
        IO::write(im, (dir / "img.den").string());
        const auto dm = IO::DenFileReader{(dir / "img.den").string()}.readMetadata();
        CHECK(dm.num_rows == 2u && dm.num_cols == 3u && dm.num_slices == 1u && dm.type == IO::ArrayElemType::T_FLOAT);
        fs::remove_all(dir);
    }
    // projection matrices (proj_geom_factory_proj_mat_test.cpp:16-45): the C-arm matrix
    // decomposed, its view made by the factory (focal length in pixels, times 0.616 mm),
    // and composed back; then two matrices through a MetaImage file and ProjMatReader
    {
        const std::array<double, 12> artis = {2.48159483828274, 0.39171974522293, 0.0, 307.5,
                                              0.0, 0.305095541401274, -2.48159483828274, 239.5,
                                              0.0, 0.00127388535031847, 0.0, 1.0};
        constexpr float pix = 0.616f;
        const ProjectionMatrix dec{artis};
        auto pg = ProjGeomFactory<DetFanLine, DetFanLine>::create(dec);
        pg.src_det_dist.val *= pix;
        const float ang = std::atan2(pg.src_pos.y.val, pg.src_pos.x.val);
        CHECK(near(ang < 0.0f ? ang + 2.0f * std::numbers::pi_v<float> : ang, 3.0f * std::numbers::pi_v<float> / 2.0f, 1e-6f));
        CHECK(near(pg.src_det_dist.val, 1200.0f, 1e-3f));
        CHECK(pg.det_row_center.val == 307.5f && pg.det_col_center.val == 239.5f);
        CHECK(near(pg.det_row_dir.x.val, 1.0f, 1e-8f) && near(pg.det_row_dir.y.val, 0.0f, 1e-8f) &&
              near(pg.det_row_dir.z.val, 0.0f, 1e-8f));
        CHECK(near(pg.det_col_dir.x.val, 0.0f, 1e-8f) && near(pg.det_col_dir.y.val, 0.0f, 1e-8f) &&
              near(pg.det_col_dir.z.val, -1.0f, 1e-8f));
        const ProjectionMatrix comp{pg, pix, pix};
        CHECK(near(dec.focalLengthDetRow(), comp.focalLengthDetRow(), 1e-2f));
        CHECK(near(dec.principalPointDetRow(), comp.principalPointDetRow()) &&
              near(dec.principalPointDetCol(), comp.principalPointDetCol()));
        const auto K = dec.intrinsicMatrix();
        CHECK(near(K[0][0], dec.focalLengthDetRow(), 1e-2f) && K[2][2] == 1.0f && K[1][0] == 0.0f);
        const auto M = comp.matrix();
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 4; ++j) CHECK(near(M[i][j], float(artis[i * 4 + j]), 1e-4f * (1.0f + std::fabs(float(artis[i * 4 + j])))));

        namespace fs = std::filesystem;
        const fs::path dir = fs::temp_directory_path() / ("gct_pmat_" + std::to_string(::getpid()));
        fs::create_directories(dir);
        std::vector<double> two(artis.begin(), artis.end());
        for (int j = 0; j < 12; ++j) two.push_back(artis[j] * 2.0);   // the same view, scaled matrix
        IO::ArrayFileWriter::MetaImgMetaData<3u> md;
        md.dim_size = {4u, 3u, 2u};
        const std::string f = (dir / "pmats.mha").string();
        IO::ArrayFileWriter::write(md, two, f);
        const auto mats = IO::ProjMatReader<IO::MetaImgFileReader>::read(f);
        CHECK(mats.size() == 2u);
        CHECK(mats.size() == 2u && near(mats[1].focalLengthDetRow(), dec.focalLengthDetRow(), 1e-2f) &&
              near(mats[1].principalPointDetRow(), 307.5f));
        md.dim_size = {11u, 1u, 2u};   // 11 elements per slice: not a projection matrix
        const std::string bad = (dir / "bad.mha").string();
        IO::ArrayFileWriter::write(md, std::vector<double>(22, 1.0), bad);
        bool threw = false;
        try {
            (void)IO::ProjMatReader<IO::MetaImgFileReader>::read(bad);
        } catch (std::runtime_error const &e) {
            threw = std::string(e.what()).find("12 elements") != std::string::npos;
        }
        CHECK(threw);
        fs::remove_all(dir);
    }

    if (argc >= 3 && std::strcmp(argv[1], "--gpu") == 0) {
        // flat-panel full scan through the drop-in call, 48^3 from a uniform sphere's exact projections
        ConeBeamProjGeomParams p;
        p.rot.init_ang = Radians{0.0f};
        p.rot.ang_incr = Radians{2.0f * std::numbers::pi_v<float> / 90.0f};
        p.rot.num_proj = 90u;
        p.det_row_center = 39.5_pix;
        p.det_col_center = 31.5_pix;
        p.src_iso_dist = 200.0_mm;
        p.src_det_dist = 400.0_mm;
        std::vector<ProjGeom<DetFanLine, DetFanLine>> v;
        for (unsigned i = 0; i < 90u; ++i) v.push_back(ProjGeomFactory<DetFanLine, DetFanLine>::create(p, i));
        AxialCTGeom<DetFanLine, DetFanLine> cg{std::make_tuple(DetFanLine{{80u, 1.0_mm}}, DetFanLine{{64u, 1.0_mm}}), v};
        Data::ProjectionStack<float, 2u> proj{{80u, 64u}, 90u};
        std::ifstream in(std::string(argv[2]) + "/proj.raw", std::ios::binary);
        std::vector<float> buf(proj.numElements());
        in.read(reinterpret_cast<char *>(buf.data()), static_cast<std::streamsize>(buf.size() * sizeof(float)));
        proj.setData(buf);
        CHECK(Recon::isValidAxialRecon(cg, proj));
        Data::ReconImg<float, 3u> vol{ReconImgGeom<3u>{{48u, 0.5_mm}, {48u, 0.5_mm}, {48u, 0.5_mm}, {}}};
        Recon::reconstruct(cg, proj, vol);
        std::ofstream out(std::string(argv[2]) + "/vol.raw", std::ios::binary);
        out.write(reinterpret_cast<const char *>(vol.elements().data()),
                  static_cast<std::streamsize>(vol.elements().size() * sizeof(float)));
    }
    std::printf("%s (%d failures)\n", failures ? "FAILED" : "OK", failures);
    return failures ? 1 : 0;
}
