// gct.hpp -- the C++ drop-in API of the MI355X FDK library.
//
// Same type names, fields and call shape as the reference's public API
// (satidev/gct: src/cpp/geom, src/cpp/data, src/cuda/recon/recon_wrappers.cuh),
// header-only over the C ABI in gct_amd.h.  A caller of
//     GCT::Recon::CUDA::reconstruct(AxialCTGeom<...>, ProjectionStack, ReconImg&)
// calls GCT::Recon::reconstruct with the same arguments; errors surface as the
// same exception types (std::invalid_argument / std::runtime_error).
#pragma once

#include "gct_amd.h"

#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <numbers>
#include <optional>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <utility>
#include <vector>

namespace GCT {

namespace Detail {
inline void check(gct_status s)
{
    if (s == GCT_OK) return;
    const std::string msg = gct_last_error();
    if (s == GCT_ERR_INVALID_ARGUMENT) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}
}  // namespace Detail

// ---- units (src/cpp/geom/units.h) --------------------------------------------
namespace Geom::Units {
struct MilliMeters { float val; };
struct DetPixels { float val; };
struct Radians { float val; };
constexpr Radians operator""_rad(long double v) { return Radians{static_cast<float>(v)}; }
constexpr Radians operator""_deg(long double v)
{
    return Radians{static_cast<float>(std::numbers::pi_v<long double> * v / 180.0)};
}
constexpr MilliMeters operator""_mm(long double v) { return MilliMeters{static_cast<float>(v)}; }
constexpr DetPixels operator""_pix(long double v) { return DetPixels{static_cast<float>(v)}; }
template <class U> constexpr bool operator==(U a, U b) requires requires { a.val; } { return a.val == b.val; }
template <class U> constexpr bool operator<(U a, U b) requires requires { a.val; } { return a.val < b.val; }
template <class U> constexpr bool operator>(U a, U b) requires requires { a.val; } { return a.val > b.val; }
}  // namespace Geom::Units

// ---- geometry (src/cpp/geom) -------------------------------------------------
namespace Geom {

template <size_t Dim> struct VecWCS;
template <> struct VecWCS<2u> { Units::MilliMeters x, y; };
template <> struct VecWCS<3u> { Units::MilliMeters x, y, z; };

template <typename WidthUnit> struct Sampler1D {
    unsigned num_samples;
    WidthUnit sampling_width;
};
using EqDistSampler = Sampler1D<Units::MilliMeters>;
using EqAngSampler = Sampler1D<Units::Radians>;

struct DetParallel { EqDistSampler sampler; };
struct DetFanArc { EqAngSampler sampler; };
struct DetFanLine { EqDistSampler sampler; };

template <typename DetRow, typename... DetCol> struct ProjGeom;
template <> struct ProjGeom<DetParallel> {
    Units::Radians rot_ang;
    Units::DetPixels det_row_center;
};
template <> struct ProjGeom<DetFanArc> {
    Units::MilliMeters src_det_dist;
    Units::DetPixels det_row_center;
    VecWCS<2u> src_pos, det_row_dir, iso_ray_dir;
};
template <> struct ProjGeom<DetFanLine> {
    Units::MilliMeters src_det_dist;
    Units::DetPixels det_row_center;
    VecWCS<2u> src_pos, det_row_dir;
};
template <> struct ProjGeom<DetParallel, DetParallel> {
    Units::Radians rot_ang;
    Units::DetPixels det_row_center, det_col_center;
};
template <> struct ProjGeom<DetParallel, DetFanLine> {
    Units::MilliMeters src_det_dist, src_iso_dist;
    Units::Radians rot_ang;
    Units::DetPixels det_row_center, det_col_center;
};
template <> struct ProjGeom<DetFanArc, DetFanLine> {
    Units::MilliMeters src_det_dist;
    Units::DetPixels det_row_center, det_col_center;
    VecWCS<3u> src_pos, det_row_dir, det_col_dir, iso_ray_dir;
};
template <> struct ProjGeom<DetFanLine, DetFanLine> {
    Units::MilliMeters src_det_dist;
    Units::DetPixels det_row_center, det_col_center;
    VecWCS<3u> src_pos, det_row_dir, det_col_dir;
};

enum class RotDir { CCW = -1, CW = 1 };

struct RotationGeom {
    RotDir dir;
    Units::Radians rot_ang_incr;
    Units::Radians rot_ang_range;
    unsigned num_proj_per_rot;
    std::vector<Units::Radians> rot_ang;
    std::vector<Units::Radians> rel_rot_ang;
};

template <typename DetRow, typename... DetCol> constexpr int32_t geomKind()
{
    if constexpr (sizeof...(DetCol) == 0) {
        if constexpr (std::is_same_v<DetRow, DetParallel>) return GCT_PARALLEL_1D;
        else if constexpr (std::is_same_v<DetRow, DetFanArc>) return GCT_FAN_ARC_1D;
        else return GCT_FAN_LINE_1D;
    } else {
        using DC = std::tuple_element_t<0, std::tuple<DetCol...>>;
        if constexpr (std::is_same_v<DetRow, DetParallel> && std::is_same_v<DC, DetParallel>) return GCT_PARALLEL_2D;
        else if constexpr (std::is_same_v<DetRow, DetParallel>) return GCT_CONE_PARALLEL;
        else if constexpr (std::is_same_v<DetRow, DetFanArc>) return GCT_CYLINDRICAL;
        else return GCT_FLAT;
    }
}

namespace Detail {
inline void put(float *d, VecWCS<2u> const &v) { d[0] = v.x.val; d[1] = v.y.val; d[2] = 0.0f; }
inline void put(float *d, VecWCS<3u> const &v) { d[0] = v.x.val; d[1] = v.y.val; d[2] = v.z.val; }
inline void get(float const *d, VecWCS<2u> &v) { v.x.val = d[0]; v.y.val = d[1]; }
inline void get(float const *d, VecWCS<3u> &v) { v.x.val = d[0]; v.y.val = d[1]; v.z.val = d[2]; }

template <typename DetRow, typename... DetCol>
gct_proj_geom toC(ProjGeom<DetRow, DetCol...> const &pg)
{
    gct_proj_geom g{};
    using P = ProjGeom<DetRow, DetCol...>;
    if constexpr (requires(P p) { p.rot_ang; }) g.rot_ang = pg.rot_ang.val;
    if constexpr (requires(P p) { p.src_det_dist; }) g.src_det_dist = pg.src_det_dist.val;
    if constexpr (requires(P p) { p.src_iso_dist; }) g.src_iso_dist = pg.src_iso_dist.val;
    g.det_row_center = pg.det_row_center.val;
    if constexpr (requires(P p) { p.det_col_center; }) g.det_col_center = pg.det_col_center.val;
    if constexpr (requires(P p) { p.src_pos; }) put(g.src_pos, pg.src_pos);
    if constexpr (requires(P p) { p.det_row_dir; }) put(g.det_row_dir, pg.det_row_dir);
    if constexpr (requires(P p) { p.det_col_dir; }) put(g.det_col_dir, pg.det_col_dir);
    if constexpr (requires(P p) { p.iso_ray_dir; }) put(g.iso_ray_dir, pg.iso_ray_dir);
    return g;
}

template <typename DetRow, typename... DetCol>
ProjGeom<DetRow, DetCol...> fromC(gct_proj_geom const &g)
{
    ProjGeom<DetRow, DetCol...> pg{};
    using P = ProjGeom<DetRow, DetCol...>;
    if constexpr (requires(P p) { p.rot_ang; }) pg.rot_ang.val = g.rot_ang;
    if constexpr (requires(P p) { p.src_det_dist; }) pg.src_det_dist.val = g.src_det_dist;
    if constexpr (requires(P p) { p.src_iso_dist; }) pg.src_iso_dist.val = g.src_iso_dist;
    pg.det_row_center.val = g.det_row_center;
    if constexpr (requires(P p) { p.det_col_center; }) pg.det_col_center.val = g.det_col_center;
    if constexpr (requires(P p) { p.src_pos; }) get(g.src_pos, pg.src_pos);
    if constexpr (requires(P p) { p.det_row_dir; }) get(g.det_row_dir, pg.det_row_dir);
    if constexpr (requires(P p) { p.det_col_dir; }) get(g.det_col_dir, pg.det_col_dir);
    if constexpr (requires(P p) { p.iso_ray_dir; }) get(g.iso_ray_dir, pg.iso_ray_dir);
    return pg;
}

template <typename Det> gct_sampler sampler(Det const &d)
{
    return gct_sampler{d.sampler.num_samples, d.sampler.sampling_width.val};
}
}  // namespace Detail

// CTGeom / AxialCTGeom (src/cpp/geom/ct_geom.h:16-53, axial_ct_geom.h:8-13):
// the constructor analyses the trajectory and throws for an invalid one.
template <typename DetRow, typename... DetCol>
struct CTGeom {
public:
    CTGeom(std::tuple<DetRow, DetCol...> const &t_dg, std::vector<ProjGeom<DetRow, DetCol...>> const &t_fpg)
        : dg{t_dg}, fpg{t_fpg}
    {
        cviews.reserve(fpg.size());
        for (auto const &p : fpg) cviews.push_back(Detail::toC(p));
        gct_ct_geom g = c();
        gct_rotation_geom r{};
        std::vector<float> rot(fpg.size()), rel(fpg.size());
        if (gct_rotation_geom_compute(&g, &r, rot.data(), rel.data()) != GCT_OK)
            throw std::runtime_error{"Invalid X-ray source trajectory.\n"};
        rg.dir = r.dir < 0 ? RotDir::CCW : RotDir::CW;
        rg.rot_ang_incr = Units::Radians{r.rot_ang_incr};
        rg.rot_ang_range = Units::Radians{r.rot_ang_range};
        rg.num_proj_per_rot = r.num_proj_per_rot;
        for (size_t i = 0; i < fpg.size(); ++i) {
            rg.rot_ang.push_back(Units::Radians{rot[i]});
            rg.rel_rot_ang.push_back(Units::Radians{rel[i]});
        }
    }
    std::tuple<DetRow, DetCol...> const &detGeom() const noexcept { return dg; }
    std::vector<ProjGeom<DetRow, DetCol...>> const &projGeom() const noexcept { return fpg; }
    RotationGeom const &rotationGeom() const noexcept { return rg; }

    // The C-ABI view of this geometry (valid while *this lives).
    gct_ct_geom c() const
    {
        gct_ct_geom g{};
        g.kind = geomKind<DetRow, DetCol...>();
        g.det_row = Detail::sampler(std::get<0>(dg));
        if constexpr (sizeof...(DetCol) > 0) g.det_col = Detail::sampler(std::get<1>(dg));
        else g.det_col = gct_sampler{1u, 1.0f};
        g.num_proj = static_cast<uint32_t>(cviews.size());
        g.proj_geom = cviews.data();
        return g;
    }

protected:
    std::tuple<DetRow, DetCol...> dg;
    std::vector<ProjGeom<DetRow, DetCol...>> fpg;
    RotationGeom rg;
    std::vector<gct_proj_geom> cviews;
};

template <typename DetRow, typename... DetCol>
struct AxialCTGeom : public CTGeom<DetRow, DetCol...> {
    using CTGeom<DetRow, DetCol...>::CTGeom;
};

// SpiralCTGeom (spiral_ct_geom.h:10-95): a 2-D detector on a helix; the
// constructor throws std::runtime_error for zero projections per rotation,
// unsorted source z or a non-positive pitch.
template <typename DetRow, typename DetCol>
struct SpiralCTGeom : public CTGeom<DetRow, DetCol> {
    SpiralCTGeom(std::tuple<DetRow, DetCol> const &t_dg, std::vector<ProjGeom<DetRow, DetCol>> const &t_fpg)
        : CTGeom<DetRow, DetCol>{t_dg, t_fpg}
    {
        // validity is the window computation's: it throws what the reference's constructor throws
        gct_ct_geom g = this->c();
        gct_image_geom probe{};
        probe.dim = 3u;
        probe.x = probe.y = probe.z = gct_sampler{1u, 1.0f};
        uint32_t w[2];
        const gct_status st = gct_spiral_windows(&g, &probe, w);
        if (st != GCT_OK) throw std::runtime_error{gct_last_error()};
    }
    // [first, last) views of the rotation centred on the slice at slice_pos_mm.
    std::pair<unsigned, unsigned> firstAndLastProjIndices(float slice_pos_mm) const noexcept
    {
        unsigned mid = 0;
        float best = std::fabs(this->fpg[0].src_pos.z.val - slice_pos_mm);
        for (unsigned v = 1; v < this->fpg.size(); ++v) {
            const float d = std::fabs(this->fpg[v].src_pos.z.val - slice_pos_mm);
            if (d < best) { best = d; mid = v; }
        }
        const unsigned first = mid - this->rg.num_proj_per_rot / 2u;
        return {first, first + this->rg.num_proj_per_rot};
    }
};

// ReconImgGeom<2|3> (src/cpp/geom/recon_img_geom.h:12-30).
template <size_t Dim> struct ReconImgGeom;
template <> struct ReconImgGeom<2u> {
    EqDistSampler x, y;
    VecWCS<2u> mid_point;
};
template <> struct ReconImgGeom<3u> {
    EqDistSampler x, y, z;
    VecWCS<3u> mid_point;
};

// Acquisition parameters (src/cpp/io/parameters/proj_geom_params.h:16-59).
enum class WCS { DCM };
struct RotationParams {
    RotDir dir = RotDir::CW;
    Units::Radians init_ang;
    Units::Radians ang_incr;
    unsigned num_proj;
    float hel_pitch = 0.0f;
};
struct Parallel1DProjGeomParams {
    Units::DetPixels det_row_center;
    WCS wcs = WCS::DCM;
    RotationParams rot;
};
struct FanBeamProjGeomParams {
    Units::DetPixels det_row_center;
    WCS wcs = WCS::DCM;
    RotationParams rot;
    Units::MilliMeters src_det_dist;
    Units::MilliMeters src_iso_dist;
};
struct Parallel2DProjGeomParams {
    Units::DetPixels det_row_center;
    WCS wcs = WCS::DCM;
    RotationParams rot;
    Units::DetPixels det_col_center;
};
struct ConeBeamProjGeomParams {
    Units::DetPixels det_row_center;
    WCS wcs = WCS::DCM;
    RotationParams rot;
    Units::MilliMeters src_det_dist;
    Units::MilliMeters src_iso_dist;
    Units::DetPixels det_col_center;
};

// ProjectionMatrix (src/cpp/geom/projection_matrix.h:14-72): a flat-panel
// view as a 3x4 matrix P = K [R | -R s] normalised by P(2,3), composed from a
// view and the pixel sizes (projection_matrix.cpp:66-96) or decomposed into
// the intrinsic matrix K (positive diagonal, K(2,2) = 1), the rotation R and
// the source position s (:4-65).  The reference holds Eigen matrices; here
// they are row-major arrays of the same FP32 values.
class ProjectionMatrix {
public:
    using Mat34 = std::array<std::array<float, 4>, 3>;
    using Mat33 = std::array<std::array<float, 3>, 3>;

    explicit ProjectionMatrix(ProjGeom<DetFanLine, DetFanLine> const &pg, float pix_width_mm, float pix_height_mm)
    {
        const gct_proj_geom g = Detail::toC(pg);
        double p[12];
        GCT::Detail::check(gct_projection_matrix(&g, pix_width_mm, pix_height_mm, p));
        set(p);
        focal_ = pg.src_det_dist.val / pix_width_mm;
        ppu_ = pg.det_row_center.val;
        ppv_ = pg.det_col_center.val;
        intr_ = {{{focal_, 0.0f, ppu_}, {0.0f, pg.src_det_dist.val / pix_height_mm, ppv_}, {0.0f, 0.0f, 1.0f}}};
        rot_ = {{{g.det_row_dir[0], g.det_row_dir[1], g.det_row_dir[2]},
                 {g.det_col_dir[0], g.det_col_dir[1], g.det_col_dir[2]},
                 {g.det_row_dir[1] * g.det_col_dir[2] - g.det_row_dir[2] * g.det_col_dir[1],
                  g.det_row_dir[2] * g.det_col_dir[0] - g.det_row_dir[0] * g.det_col_dir[2],
                  g.det_row_dir[0] * g.det_col_dir[1] - g.det_row_dir[1] * g.det_col_dir[0]}}};
        src_ = {g.src_pos[0], g.src_pos[1], g.src_pos[2]};
    }
    explicit ProjectionMatrix(std::array<double, 12u> const &mat)
    {
        gct_proj_geom g{};
        double f[2];
        GCT::Detail::check(gct_projection_matrix_decompose(mat.data(), 1.0f, &g, f));
        double p[12];
        std::copy(mat.begin(), mat.end(), p);
        set(p);
        focal_ = static_cast<float>(f[0]);
        ppu_ = g.det_row_center;
        ppv_ = g.det_col_center;
        rot_ = {{{g.det_row_dir[0], g.det_row_dir[1], g.det_row_dir[2]},
                 {g.det_col_dir[0], g.det_col_dir[1], g.det_col_dir[2]},
                 {g.iso_ray_dir[0], g.iso_ray_dir[1], g.iso_ray_dir[2]}}};
        src_ = {g.src_pos[0], g.src_pos[1], g.src_pos[2]};
        // K = M R^T (M the left 3x3 of P), normalised so that K(2,2) = 1
        double k[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                k[i][j] = mat[i * 4 + 0] * rot_[j][0] + mat[i * 4 + 1] * rot_[j][1] + mat[i * 4 + 2] * rot_[j][2];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) intr_[i][j] = static_cast<float>(j < i ? 0.0 : k[i][j] / k[2][2]);
    }

    Mat34 matrix() const noexcept { return mat_; }
    float principalPointDetRow() const noexcept { return ppu_; }
    float principalPointDetCol() const noexcept { return ppv_; }
    // focal length (source-detector distance) in pixels of a detector row
    float focalLengthDetRow() const noexcept { return focal_; }
    std::array<float, 3> srcPos() const noexcept { return src_; }
    Mat33 rotationMatrix() const noexcept { return rot_; }
    Mat33 intrinsicMatrix() const noexcept { return intr_; }

private:
    void set(const double *p)
    {
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 4; ++j) mat_[i][j] = static_cast<float>(p[i * 4 + j]);
    }
    Mat34 mat_{};
    Mat33 intr_{}, rot_{};
    float ppu_ = 0.0f, ppv_ = 0.0f, focal_ = 0.0f;
    std::array<float, 3> src_{};
};

// ProjGeomFactory<...>::create (src/cpp/io/geom_reader/proj_geom_factory_params.cpp:8-108).
template <typename DetRow, typename... DetCol>
struct ProjGeomFactory {
    // create(ProjectionMatrix) (proj_geom_factory_proj_mat.cpp:7-26), flat panels
    // only: as in the reference the source-detector distance is the focal
    // length in pixels of a detector row -- the caller multiplies it by the
    // pixel width (the reference's TODO at :13)
    static ProjGeom<DetRow, DetCol...> create(ProjectionMatrix const &m)
        requires(std::is_same_v<ProjGeom<DetRow, DetCol...>, ProjGeom<DetFanLine, DetFanLine>>)
    {
        ProjGeom<DetFanLine, DetFanLine> pg{};
        pg.src_det_dist.val = m.focalLengthDetRow();
        pg.det_row_center.val = m.principalPointDetRow();
        pg.det_col_center.val = m.principalPointDetCol();
        const auto s = m.srcPos();
        const auto r = m.rotationMatrix();
        pg.src_pos = VecWCS<3u>{{s[0]}, {s[1]}, {s[2]}};
        pg.det_row_dir = VecWCS<3u>{{r[0][0]}, {r[0][1]}, {r[0][2]}};
        pg.det_col_dir = VecWCS<3u>{{r[1][0]}, {r[1][1]}, {r[1][2]}};
        return pg;
    }

    template <typename Params>
    static ProjGeom<DetRow, DetCol...> create(Params const &p, unsigned proj_idx)
    {
        gct_circular_params c{};
        c.kind = geomKind<DetRow, DetCol...>();
        c.dir = static_cast<int32_t>(p.rot.dir);
        c.init_ang = p.rot.init_ang.val;
        c.ang_incr = p.rot.ang_incr.val;
        c.num_proj = proj_idx + 1u;
        c.hel_pitch = p.rot.hel_pitch;
        c.det_row_center = p.det_row_center.val;
        if constexpr (requires(Params q) { q.det_col_center; }) c.det_col_center = p.det_col_center.val;
        if constexpr (requires(Params q) { q.src_det_dist; }) c.src_det_dist = p.src_det_dist.val;
        if constexpr (requires(Params q) { q.src_iso_dist; }) c.src_iso_dist = p.src_iso_dist.val;
        std::vector<gct_proj_geom> out(c.num_proj);
        GCT::Detail::check(gct_geom_create_circular(&c, out.data()));
        return Detail::fromC<DetRow, DetCol...>(out[proj_idx]);
    }
};

}  // namespace Geom

// ---- data containers (src/cpp/data) ---------------------------------------------
namespace Data {

template <size_t ProjDim> struct ProjSize;
template <> struct ProjSize<1u> { unsigned num_cols; };
template <> struct ProjSize<2u> { unsigned num_cols, num_rows; };

template <typename... DetCol> constexpr size_t projDim() { return sizeof...(DetCol) + 1u; }
template <typename... DetCol> constexpr size_t reconImgDim() { return sizeof...(DetCol) + 2u; }

inline unsigned numProjPixels(ProjSize<1u> const &s) { return s.num_cols; }
inline unsigned numProjPixels(ProjSize<2u> const &s) { return s.num_cols * s.num_rows; }
inline unsigned numProjRows(ProjSize<1u> const &) { return 1u; }
inline unsigned numProjRows(ProjSize<2u> const &s) { return s.num_rows; }

// x (detector column) fastest, then row, then view (projection_stack.h:14).
template <typename PixType, size_t ProjDim>
class ProjectionStack {
public:
    ProjectionStack(ProjSize<ProjDim> const &size, unsigned num_proj, PixType init_val = PixType{})
        : proj_size_{size}, num_proj_{num_proj}
    {
        if (numProjPixels(size) == 0 || num_proj == 0)
            throw std::invalid_argument{"Projections size or number of projections cannot be zero.\n"};
        elems_.assign(size_t(numProjPixels(size)) * num_proj, init_val);
    }
    ProjSize<ProjDim> projSize() const noexcept { return proj_size_; }
    unsigned numProj() const noexcept { return num_proj_; }
    std::vector<PixType> const &elements() const { return elems_; }
    std::vector<PixType> &elements() { return elems_; }
    size_t numElements() const noexcept { return elems_.size(); }
    void setData(std::vector<PixType> data)
    {
        if (data.size() != elems_.size())
            throw std::invalid_argument{"Cannot copy data to the projection stack: wrong number of elements.\n"};
        elems_ = std::move(data);
    }

private:
    ProjSize<ProjDim> proj_size_;
    unsigned num_proj_;
    std::vector<PixType> elems_;
};

// x fastest, then y, then z (recon_img.h, back_proj.cuh:76).
template <typename PixType, size_t Dim>
class ReconImg {
public:
    explicit ReconImg(Geom::ReconImgGeom<Dim> const &geom) : geom_{geom}
    {
        size_t n = size_t(geom.x.num_samples) * geom.y.num_samples;
        if constexpr (Dim == 3u) n *= geom.z.num_samples;
        bool ok = geom.x.sampling_width.val > 0.0f && geom.y.sampling_width.val > 0.0f;
        if constexpr (Dim == 3u) ok = ok && geom.z.sampling_width.val > 0.0f;
        if (n == 0) throw std::invalid_argument{"The number of samples in any dimension cannot be zero.\n"};
        if (!ok) throw std::invalid_argument{"The sampling width of the image in any dimension cannot be zero or negative.\n"};
        elems_.assign(n, PixType{});
    }
    Geom::ReconImgGeom<Dim> const &geom() const noexcept { return geom_; }
    std::vector<PixType> const &elements() const { return elems_; }
    size_t numElements() const noexcept { return elems_.size(); }
    void setData(std::vector<PixType> &&data)
    {
        if (data.size() != elems_.size())
            throw std::invalid_argument{"Cannot move data to the recon image: wrong number of elements.\n"};
        elems_ = std::move(data);
    }
    gct_image_geom c() const
    {
        gct_image_geom g{};
        g.dim = static_cast<uint32_t>(Dim);
        g.x = gct_sampler{geom_.x.num_samples, geom_.x.sampling_width.val};
        g.y = gct_sampler{geom_.y.num_samples, geom_.y.sampling_width.val};
        g.mid_point[0] = geom_.mid_point.x.val;
        g.mid_point[1] = geom_.mid_point.y.val;
        if constexpr (Dim == 3u) {
            g.z = gct_sampler{geom_.z.num_samples, geom_.z.sampling_width.val};
            g.mid_point[2] = geom_.mid_point.z.val;
        } else {
            g.z = gct_sampler{1u, 1.0f};
        }
        return g;
    }

private:
    Geom::ReconImgGeom<Dim> geom_;
    std::vector<PixType> elems_;
};

}  // namespace Data

// ---- reconstruction (src/cuda/recon/recon_wrappers.cuh:13-45) -----------------
namespace Recon {

enum class ApodFunc { NONE, HANN, HAMMING };
struct FiltParams {
    float cut_off_freq = 1.0f;
    ApodFunc apod_func = ApodFunc::NONE;
};

// isValidAxialRecon (src/cpp/recon/check_validity.h:66-75).
template <typename DetRow, typename... DetCol>
bool isValidAxialRecon(Geom::CTGeom<DetRow, DetCol...> const &cg,
                       Data::ProjectionStack<float, Data::projDim<DetCol...>()> const &proj, bool verbose = false)
{
    gct_ct_geom g = cg.c();
    const auto ps = proj.projSize();
    uint32_t rows = 1u;
    if constexpr (sizeof...(DetCol) > 0) rows = ps.num_rows;
    const gct_status s = gct_check_axial_recon(&g, ps.num_cols, rows, proj.numProj());
    (void)verbose;
    return s == GCT_OK;
}

namespace Detail {
template <typename DetRow, typename... DetCol>
void reconstructImpl(Geom::AxialCTGeom<DetRow, DetCol...> const &cg,
                     Data::ProjectionStack<float, Data::projDim<DetCol...>()> const &proj,
                     Data::ReconImg<float, Data::reconImgDim<DetCol...>()> &vol, gct_recon_opts const *opts)
{
    gct_ct_geom g = cg.c();
    gct_image_geom im = vol.c();
    std::vector<float> out(vol.numElements());
    GCT::Detail::check(gct_reconstruct(&g, proj.elements().data(), &im, out.data(), opts));
    vol.setData(std::move(out));   // the result replaces the image's data (fdk_axial.cuh:202-204)
}
}  // namespace Detail

// The seven overloads of Recon::CUDA::reconstruct.
inline void reconstruct(Geom::AxialCTGeom<Geom::DetParallel> const &cg, Data::ProjectionStack<float, 1u> const &proj,
                        Data::ReconImg<float, 2u> &vol, gct_recon_opts const *opts = nullptr)
{ Detail::reconstructImpl(cg, proj, vol, opts); }
inline void reconstruct(Geom::AxialCTGeom<Geom::DetFanArc> const &cg, Data::ProjectionStack<float, 1u> const &proj,
                        Data::ReconImg<float, 2u> &vol, gct_recon_opts const *opts = nullptr)
{ Detail::reconstructImpl(cg, proj, vol, opts); }
inline void reconstruct(Geom::AxialCTGeom<Geom::DetFanLine> const &cg, Data::ProjectionStack<float, 1u> const &proj,
                        Data::ReconImg<float, 2u> &vol, gct_recon_opts const *opts = nullptr)
{ Detail::reconstructImpl(cg, proj, vol, opts); }
inline void reconstruct(Geom::AxialCTGeom<Geom::DetParallel, Geom::DetParallel> const &cg,
                        Data::ProjectionStack<float, 2u> const &proj, Data::ReconImg<float, 3u> &vol,
                        gct_recon_opts const *opts = nullptr)
{ Detail::reconstructImpl(cg, proj, vol, opts); }
inline void reconstruct(Geom::AxialCTGeom<Geom::DetParallel, Geom::DetFanLine> const &cg,
                        Data::ProjectionStack<float, 2u> const &proj, Data::ReconImg<float, 3u> &vol,
                        gct_recon_opts const *opts = nullptr)
{ Detail::reconstructImpl(cg, proj, vol, opts); }
inline void reconstruct(Geom::AxialCTGeom<Geom::DetFanArc, Geom::DetFanLine> const &cg,
                        Data::ProjectionStack<float, 2u> const &proj, Data::ReconImg<float, 3u> &vol,
                        gct_recon_opts const *opts = nullptr)
{ Detail::reconstructImpl(cg, proj, vol, opts); }
inline void reconstruct(Geom::AxialCTGeom<Geom::DetFanLine, Geom::DetFanLine> const &cg,
                        Data::ProjectionStack<float, 2u> const &proj, Data::ReconImg<float, 3u> &vol,
                        gct_recon_opts const *opts = nullptr)
{ Detail::reconstructImpl(cg, proj, vol, opts); }

// Spiral FDK (recon_wrappers.cu:58-63, fdk_spiral.cuh:28-94); the flat-panel
// overload is an extension.
inline void reconstruct(Geom::SpiralCTGeom<Geom::DetFanArc, Geom::DetFanLine> const &cg,
                        Data::ProjectionStack<float, 2u> const &proj, Data::ReconImg<float, 3u> &vol,
                        gct_recon_opts const *opts = nullptr)
{
    gct_ct_geom g = cg.c();
    gct_image_geom im = vol.c();
    std::vector<float> out(vol.numElements());
    GCT::Detail::check(gct_reconstruct_spiral(&g, proj.elements().data(), &im, out.data(), opts));
    vol.setData(std::move(out));
}
inline void reconstruct(Geom::SpiralCTGeom<Geom::DetFanLine, Geom::DetFanLine> const &cg,
                        Data::ProjectionStack<float, 2u> const &proj, Data::ReconImg<float, 3u> &vol,
                        gct_recon_opts const *opts = nullptr)
{
    gct_ct_geom g = cg.c();
    gct_image_geom im = vol.c();
    std::vector<float> out(vol.numElements());
    GCT::Detail::check(gct_reconstruct_spiral(&g, proj.elements().data(), &im, out.data(), opts));
    vol.setData(std::move(out));
}

}  // namespace Recon
// ---- array files (src/cpp/io): readers, writers, projection and image I/O -------
namespace IO {

enum class ArrayElemType { T_UINT8, T_INT8, T_UINT16, T_INT16, T_UINT32, T_INT32, T_UINT64, T_INT64, T_FLOAT, T_DOUBLE };

struct ArrayMetadata {
    unsigned num_rows;
    unsigned num_cols;
    unsigned num_slices;   // matrices per file
    ArrayElemType type;
};

namespace Detail {
template <typename T> constexpr int32_t elemType()
{
    if constexpr (std::is_same_v<T, std::uint8_t>) return GCT_T_UINT8;
    else if constexpr (std::is_same_v<T, std::int8_t>) return GCT_T_INT8;
    else if constexpr (std::is_same_v<T, std::uint16_t>) return GCT_T_UINT16;
    else if constexpr (std::is_same_v<T, std::int16_t>) return GCT_T_INT16;
    else if constexpr (std::is_same_v<T, std::uint32_t>) return GCT_T_UINT32;
    else if constexpr (std::is_same_v<T, std::int32_t>) return GCT_T_INT32;
    else if constexpr (std::is_same_v<T, std::uint64_t>) return GCT_T_UINT64;
    else if constexpr (std::is_same_v<T, std::int64_t>) return GCT_T_INT64;
    else if constexpr (std::is_same_v<T, float>) return GCT_T_FLOAT;
    else {
        static_assert(std::is_same_v<T, double>, "unsupported element type");
        return GCT_T_DOUBLE;
    }
}
inline gct_array_metadata toC(ArrayMetadata const &m)
{
    return gct_array_metadata{m.num_rows, m.num_cols, m.num_slices, static_cast<int32_t>(m.type)};
}
inline ArrayMetadata fromC(gct_array_metadata const &m)
{
    return ArrayMetadata{m.num_rows, m.num_cols, m.num_slices, static_cast<ArrayElemType>(m.type)};
}
}  // namespace Detail

// ArrayFileReader (array_file_reader.h:13-72) and its three file kinds.
class ArrayFileReader {
public:
    ArrayMetadata readMetadata() const
    {
        gct_array_metadata md{};
        GCT::Detail::check(gct_file_read_metadata(filename_.c_str(), format_, rawPtr(), &md));
        return Detail::fromC(md);
    }
    // Elements of the file (or of the series), converted by static_cast.
    template <typename OpType>
    std::vector<OpType> readData() const
    {
        uint64_t n = 0;
        const int32_t t = Detail::elemType<OpType>();
        GCT::Detail::check(gct_file_read(filename_.c_str(), format_, series_, rawPtr(), t, nullptr, 0, &n));
        std::vector<OpType> out(n);
        GCT::Detail::check(gct_file_read(filename_.c_str(), format_, series_, rawPtr(), t, out.data(), n, &n));
        return out;
    }

protected:
    ArrayFileReader(std::string filename, bool series, int32_t format)
        : filename_{std::move(filename)}, series_{series ? 1 : 0}, format_{format}
    {
        uint64_t n = 0;   // existence and header checks happen on construction, as in the reference
        if (format_ != GCT_FILE_RAW)
            GCT::Detail::check(gct_file_read(filename_.c_str(), format_, series_, nullptr, GCT_T_UINT8, nullptr, 0, &n));
    }
    virtual const gct_array_metadata *rawPtr() const { return nullptr; }
    virtual ~ArrayFileReader() = default;

    std::string filename_;
    int series_;
    int32_t format_;
};

// MetaImgFileReader (meta_img_reader.h:12-47): .mha (data after the header) or .mhd + data file.
class MetaImgFileReader : public ArrayFileReader {
public:
    explicit MetaImgFileReader(std::string filename, bool series = false)
        : ArrayFileReader(std::move(filename), series, GCT_FILE_META) {}
};

// DenFileReader: three 16-bit sizes (rows, cols, slices), then the data.
class DenFileReader : public ArrayFileReader {
public:
    explicit DenFileReader(std::string filename, bool series = false)
        : ArrayFileReader(std::move(filename), series, GCT_FILE_DEN) {}
};

// RawDataFileReader (raw_data_file_reader.h:13-33): headerless, caller's metadata.
class RawDataFileReader : public ArrayFileReader {
public:
    RawDataFileReader(std::string filename, ArrayMetadata md, bool series = false)
        : ArrayFileReader(std::move(filename), series, GCT_FILE_RAW), md_{Detail::toC(md)}
    {
        uint64_t n = 0;
        GCT::Detail::check(gct_file_read(filename_.c_str(), format_, series_, &md_, GCT_T_UINT8, nullptr, 0, &n));
    }

protected:
    const gct_array_metadata *rawPtr() const override { return &md_; }

private:
    gct_array_metadata md_;
};

namespace ArrayFileWriter {

template <size_t ImgDim>
struct MetaImgMetaData {
    std::array<unsigned, ImgDim> dim_size{};
    std::array<float, ImgDim> spacing{};
    std::array<float, ImgDim> offset{};
};

struct DenMetaData {
    unsigned short num_rows;
    unsigned short num_cols;
    unsigned short num_slices;
};

namespace Detail::MetaImage {
template <typename DataElemType> std::string typeStr()
{
    return gct_meta_type_name(IO::Detail::elemType<DataElemType>());
}
template <size_t ImgDim> size_t numDataElems(std::array<unsigned, ImgDim> const &dim_size)
{
    size_t n = 1;
    for (unsigned d : dim_size) n *= d;
    return n;
}
}  // namespace Detail::MetaImage

// MetaImage writer (meta_img_file_writer.h:101-152).
template <size_t ImgDim, typename Container>
void write(MetaImgMetaData<ImgDim> const &md, Container const &data, std::string const &filename)
{
    using T = typename Container::value_type;
    GCT::Detail::check(gct_meta_write(filename.c_str(), static_cast<uint32_t>(ImgDim), md.dim_size.data(),
                                      md.spacing.data(), md.offset.data(), IO::Detail::elemType<T>(), data.data(),
                                      data.size()));
}

// Dennerlein writer (den_file_writer.h:14-60).
template <typename Container>
void write(DenMetaData const &md, Container const &data, std::string const &filename)
{
    using T = typename Container::value_type;
    GCT::Detail::check(gct_den_write(filename.c_str(), md.num_rows, md.num_cols, md.num_slices,
                                     IO::Detail::elemType<T>(), data.data(), data.size()));
}

// Headerless writer (raw_data_file_writer.h:13-36).
template <typename Container>
void write(Container const &data, std::string const &filename)
{
    using T = typename Container::value_type;
    GCT::Detail::check(gct_raw_write(filename.c_str(), IO::Detail::elemType<T>(), data.data(), data.size()));
}

}  // namespace ArrayFileWriter

// ProjectionReader (projection_reader.h:18-36): 1-D projections are the
// columns, rows x slices views; 2-D projections are cols x rows, slices views
// (projection_reader.cpp:5-17).
template <typename FileReader>
struct ProjectionReader {
    template <typename PixType, size_t Dim>
    static Data::ProjectionStack<PixType, Dim> read(std::string const &filename, bool series = false)
    {
        const FileReader reader{filename, series};
        const ArrayMetadata md = reader.readMetadata();
        std::vector<PixType> pix = reader.template readData<PixType>();
        if constexpr (Dim == 1u) {
            const unsigned views = static_cast<unsigned>(pix.size() / md.num_cols);
            Data::ProjectionStack<PixType, 1u> p{Data::ProjSize<1u>{md.num_cols}, views};
            p.setData(std::move(pix));
            return p;
        } else {
            const unsigned views = static_cast<unsigned>(pix.size() / (size_t(md.num_cols) * md.num_rows));
            Data::ProjectionStack<PixType, 2u> p{Data::ProjSize<2u>{md.num_cols, md.num_rows}, views};
            p.setData(std::move(pix));
            return p;
        }
    }
};

// ProjMatReader (proj_mat_reader.h:9-45): one 3x4 projection matrix (12
// elements, row-major, read as double) per slice of the file or series.
template <typename FileReader>
struct ProjMatReader {
    static std::vector<Geom::ProjectionMatrix> read(std::string const &filename, bool series = false)
    {
        const FileReader reader{filename, series};
        const ArrayMetadata md = reader.readMetadata();
        const std::vector<double> all = reader.template readData<double>();
        constexpr size_t kElems = 12u;
        const size_t num_mats = md.num_slices;
        if (all.size() != kElems * num_mats)
            throw std::runtime_error{"Error during the reading of projection matrix file " + filename + ".\n" +
                                     std::string{"There should be 12 elements in each projection matrix.\n"}};
        std::vector<Geom::ProjectionMatrix> out;
        out.reserve(num_mats);
        std::array<double, kElems> m{};
        for (size_t i = 0; i < num_mats; ++i) {
            std::copy_n(all.begin() + static_cast<std::ptrdiff_t>(i * kElems), kElems, m.begin());
            out.emplace_back(m);
        }
        return out;
    }
};

// Image writer (img_writer.h:20-42, img_writer.cpp:5-62): .den, or MetaImage
// with the voxel spacing and the first voxel's centre as the offset.
template <typename PixType, size_t Dim>
void write(Data::ReconImg<PixType, Dim> const &img, std::string const &filename)
{
    std::string ext;
    if (const auto dot = filename.find_last_of('.'); dot != std::string::npos) ext = filename.substr(dot);
    std::transform(ext.begin(), ext.end(), ext.begin(), [](unsigned char c) { return char(std::tolower(c)); });
    const auto &g = img.geom();
    if (ext == ".den") {
        ArrayFileWriter::DenMetaData md{static_cast<unsigned short>(g.y.num_samples),
                                        static_cast<unsigned short>(g.x.num_samples), 1u};
        if constexpr (Dim == 3u) md.num_slices = static_cast<unsigned short>(g.z.num_samples);
        ArrayFileWriter::write(md, img.elements(), filename);
    } else if (ext == ".mha" || ext == ".mhd") {
        auto first = [](Geom::EqDistSampler const &s, float mid) {
            return mid - (static_cast<float>(s.num_samples) - 1.0f) * s.sampling_width.val / 2.0f;
        };
        ArrayFileWriter::MetaImgMetaData<Dim> md;
        md.dim_size[0] = g.x.num_samples;
        md.dim_size[1] = g.y.num_samples;
        md.spacing[0] = g.x.sampling_width.val;
        md.spacing[1] = g.y.sampling_width.val;
        md.offset[0] = first(g.x, g.mid_point.x.val);
        md.offset[1] = first(g.y, g.mid_point.y.val);
        if constexpr (Dim == 3u) {
            md.dim_size[2] = g.z.num_samples;
            md.spacing[2] = g.z.sampling_width.val;
            md.offset[2] = first(g.z, g.mid_point.z.val);
        }
        ArrayFileWriter::write(md, img.elements(), filename);
    } else {
        std::fprintf(stderr, "Currently images can be saved to den or meta image file.\n");
    }
}

}  // namespace IO

}  // namespace GCT
