// C-ABI exports of reference functions that compile from the reference's own
// sources with no Eigen and no CUDA -- TEST INFRASTRUCTURE ONLY.
//
// Built by `make -C oracle ref` against the sources under the reference tree;
// the output goes to oracle/_ref/ and is used only to pin the CPU oracle
// (tests/test_oracle_ref.py).  Nothing here is shipped or called by the product.
//
// Exports:
//   ref_ramp_taps      -> GCT::Recon::projFiltKernel (src/cpp/recon/proj_filt_kernel.cpp:5-45)
//   ref_rotation_geom  -> GCT::Geom::getRotationGeom (src/cpp/geom/rotation_geom.h:90-115)
//   ref_wrap/quadrant  -> GCT::Geom::Utils (src/cpp/geom/utils.h:338-389)
#include "recon/proj_filt_kernel.h"
#include "geom/rotation_geom.h"
#include "geom/detector_1d.h"

#include <cstring>
#include <vector>

extern "C" {

// det_kind: 0 = DetParallel, 1 = DetFanLine, 2 = DetFanArc; window: 0 none, 1 Hann, 2 Hamming.
int ref_ramp_taps(int det_kind, unsigned n, float width, float cutoff, int window, float *out)
{
    GCT::Recon::FiltParams p{};
    p.cut_off_freq = cutoff;
    p.apod_func = (window == 1) ? GCT::Recon::ApodFunc::HANN
                : (window == 2) ? GCT::Recon::ApodFunc::HAMMING : GCT::Recon::ApodFunc::NONE;
    std::vector<float> taps;
    if (det_kind == 0) {
        GCT::Geom::DetParallel d{{n, GCT::Geom::Units::MilliMeters{width}}};
        taps = GCT::Recon::projFiltKernel(d, p);
    } else if (det_kind == 1) {
        GCT::Geom::DetFanLine d{{n, GCT::Geom::Units::MilliMeters{width}}};
        taps = GCT::Recon::projFiltKernel(d, p);
    } else {
        GCT::Geom::DetFanArc d{{n, GCT::Geom::Units::Radians{width}}};
        taps = GCT::Recon::projFiltKernel(d, p);
    }
    std::memcpy(out, taps.data(), taps.size() * sizeof(float));
    return static_cast<int>(taps.size());
}

// Rotation analysis of a parallel-beam angle list (rot_ang) or of fan-beam
// source positions (src_xy interleaved x,y).  Returns 0 on success, -1 when the
// reference returns std::nullopt.
int ref_rotation_geom(int fan, unsigned n, const float *rot_ang, const float *src_xy,
                      int *dir, float *incr, float *range, unsigned *nppr,
                      float *rot_out, float *rel_out)
{
    std::optional<GCT::Geom::RotationGeom> info;
    if (!fan) {
        std::vector<GCT::Geom::ProjGeom<GCT::Geom::DetParallel>> pg(n);
        for (unsigned i = 0; i < n; ++i) pg[i].rot_ang = GCT::Geom::Units::Radians{rot_ang[i]};
        info = GCT::Geom::getRotationGeom(pg);
    } else {
        std::vector<GCT::Geom::ProjGeom<GCT::Geom::DetFanLine>> pg(n);
        for (unsigned i = 0; i < n; ++i) {
            pg[i].src_pos.x.val = src_xy[2 * i];
            pg[i].src_pos.y.val = src_xy[2 * i + 1];
        }
        info = GCT::Geom::getRotationGeom(pg);
    }
    if (!info.has_value()) return -1;
    *dir = static_cast<int>(info->dir);
    *incr = info->rot_ang_incr.val;
    *range = info->rot_ang_range.val;
    *nppr = info->num_proj_per_rot;
    for (unsigned i = 0; i < n; ++i) {
        rot_out[i] = info->rot_ang[i].val;
        rel_out[i] = info->rel_rot_ang[i].val;
    }
    return 0;
}

float ref_wrap(float a) { return GCT::Geom::Utils::wrapFrom0To2PI(GCT::Geom::Units::Radians{a}).val; }
unsigned ref_quadrant(float a) { return GCT::Geom::Utils::quadrantIndex(GCT::Geom::Units::Radians{a}); }

}  // extern "C"
