// Array files of the reference's src/cpp/io: MetaImage (.mha with the data
// after the text header, .mhd with a separate data file), Dennerlein (.den:
// three 16-bit sizes, then the data) and headerless raw arrays.  Readers
// (optionally over a series of same-sized files), element-type conversion and
// writers.  Host code, no device work.
//
// Behaviour followed: MetaImgFileReader (meta_img_reader.cpp:10-165),
// DenFileReader (den_file_reader.cpp:7-61), RawDataFileReader
// (raw_data_file_reader.h:13-33), ArrayFileReader::readData / convertData /
// readDataFromBinFile (array_file_reader.h:66-150, array_file_reader.cpp:6-108),
// getAllFiles / isWritable (file_utils.cpp:7-28, file_utils.h:20-30), the
// writers (meta_img_file_writer.h:61-152, meta_img_file_writer.cpp:5-71,
// den_file_writer.h:14-60, raw_data_file_writer.h:13-36).
#include "common.hpp"

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

namespace gct {
namespace {

namespace fs = std::filesystem;

size_t elem_size(int t)
{
    switch (t) {
    case GCT_T_UINT8: case GCT_T_INT8: return 1;
    case GCT_T_UINT16: case GCT_T_INT16: return 2;
    case GCT_T_UINT32: case GCT_T_INT32: case GCT_T_FLOAT: return 4;
    case GCT_T_UINT64: case GCT_T_INT64: case GCT_T_DOUBLE: return 8;
    default: fail(GCT_ERR_INVALID_ARGUMENT, "unknown element type " + std::to_string(t));
    }
}

struct MetaName {
    const char *name;
    int type;
};
constexpr MetaName kMetaNames[] = {{"MET_UCHAR", GCT_T_UINT8},       {"MET_CHAR", GCT_T_INT8},
                                   {"MET_USHORT", GCT_T_UINT16},     {"MET_SHORT", GCT_T_INT16},
                                   {"MET_UINT", GCT_T_UINT32},       {"MET_INT", GCT_T_INT32},
                                   {"MET_ULONG_LONG", GCT_T_UINT64}, {"MET_LONG_LONG", GCT_T_INT64},
                                   {"MET_FLOAT", GCT_T_FLOAT},       {"MET_DOUBLE", GCT_T_DOUBLE}};

std::string trim(const std::string &s)
{
    const auto b = s.find_first_not_of(" \t\r\n\f\v");
    if (b == std::string::npos) return {};
    const auto e = s.find_last_not_of(" \t\r\n\f\v");
    return s.substr(b, e - b + 1);
}

std::string lower(std::string s)
{
    std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return char(std::tolower(c)); });
    return s;
}

int format_of(const std::string &filename, int format)
{
    if (format != GCT_FILE_AUTO) return format;
    const std::string ext = lower(fs::path(filename).extension().string());
    if (ext == ".mha" || ext == ".mhd") return GCT_FILE_META;
    if (ext == ".den") return GCT_FILE_DEN;
    return GCT_FILE_RAW;
}

uint64_t md_elems(const gct_array_metadata &md)
{
    return uint64_t(md.num_rows) * md.num_cols * md.num_slices;
}

void require_file(const std::string &f)
{
    if (!fs::exists(fs::path(f))) fail(GCT_ERR_RUNTIME, "File " + f + " is not found.");
}

// MetaImage header: DimSize = cols rows [slices], ElementType = MET_*; lines
// are read up to "ElementDataFile = LOCAL" (meta_img_reader.cpp:33-79).
gct_array_metadata meta_header(const std::string &fn)
{
    std::ifstream ifs(fn);
    if (!ifs) fail(GCT_ERR_RUNTIME, "File " + fn + " cannot be opened.");
    gct_array_metadata md{1u, 1u, 1u, GCT_T_UINT8};
    std::string line;
    while (std::getline(ifs, line)) {
        if (line.find("ElementDataFile = LOCAL") != std::string::npos) break;
        if (line.find("DimSize ") != std::string::npos) {
            std::stringstream ss(line.substr(line.find('=') + 1));
            std::vector<long> dims;
            std::string w;
            while (ss >> w) dims.push_back(std::stol(w));
            if (dims.size() < 2 || dims.size() > 3)
                fail(GCT_ERR_RUNTIME, "Error during reading meta image file " + fn + ": DimSize needs 2 or 3 values");
            md.num_cols = uint32_t(dims[0]);
            md.num_rows = uint32_t(dims[1]);
            if (dims.size() == 3) md.num_slices = uint32_t(dims[2]);
        }
        if (line.find("ElementType ") != std::string::npos) {
            const std::string v = trim(line.substr(line.find('=') + 1));
            bool found = false;
            for (const MetaName &m : kMetaNames)
                if (v == m.name) { md.type = m.type; found = true; }
            if (!found)
                fail(GCT_ERR_RUNTIME, "Error during reading meta image file " + fn + "\n" + v + " is not a valid ElementType");
        }
    }
    return md;
}

// Value of the first ElementDataFile tag (meta_img_reader.cpp:125-144).
std::string meta_data_tag(const std::string &fn)
{
    std::ifstream ifs(fn);
    std::string line;
    while (std::getline(ifs, line))
        if (line.find("ElementDataFile") != std::string::npos) return trim(line.substr(line.find('=') + 1));
    return {};
}

// The file holding the elements: the header itself ("LOCAL") or a file named
// by the tag in the header's directory (meta_img_reader.cpp:10-29).
std::string meta_data_file(const std::string &fn)
{
    const std::string tag = meta_data_tag(fn);
    if (tag == "LOCAL") return fn;
    const fs::path raw = fs::path(fn).parent_path() / tag;
    if (!fs::exists(raw))
        fail(GCT_ERR_RUNTIME, "Data filename " + raw.string() +
                                  " is not found.\nHeader and data files should be in the same directory.");
    return raw.string();
}

gct_array_metadata den_header(const std::string &fn)
{
    std::ifstream f(fn, std::ios::binary);
    if (!f) fail(GCT_ERR_RUNTIME, "File " + fn + " cannot be opened.");
    uint16_t h[3] = {0, 0, 0};
    f.read(reinterpret_cast<char *>(h), sizeof(h));
    const uint64_t n = uint64_t(h[0]) * h[1] * h[2];
    if (!f || n == 0) fail(GCT_ERR_RUNTIME, "Error during reading Dennerlein file " + fn + ": empty header");
    const uint64_t bytes = fs::file_size(fs::path(fn)) - sizeof(h);
    gct_array_metadata md{h[0], h[1], h[2], GCT_T_UINT8};
    // element type from the element size (den_file_reader.cpp:29-55)
    switch (bytes / n) {
    case 2: md.type = GCT_T_UINT16; break;
    case 4: md.type = GCT_T_FLOAT; break;
    case 8: md.type = GCT_T_DOUBLE; break;
    default: md.type = GCT_T_UINT8; break;
    }
    return md;
}

gct_array_metadata read_md(const std::string &fn, int format, const gct_array_metadata *raw_md)
{
    require_file(fn);
    switch (format) {
    case GCT_FILE_META: return meta_header(fn);
    case GCT_FILE_DEN: return den_header(fn);
    case GCT_FILE_RAW:
        if (!raw_md) fail(GCT_ERR_INVALID_ARGUMENT, "raw files need the caller's metadata");
        (void)elem_size(raw_md->type);
        return *raw_md;
    default: fail(GCT_ERR_INVALID_ARGUMENT, "unknown file format");
    }
}

// Bytes after `skip`; the remaining size must match exactly
// (array_file_reader.cpp:70-108).
std::vector<uint8_t> read_exact(const std::string &fn, uint64_t bytes, uint64_t skip)
{
    std::ifstream f(fn, std::ios::binary);
    if (!f) fail(GCT_ERR_RUNTIME, "File " + fn + " cannot be opened for reading data.");
    const uint64_t size = fs::file_size(fs::path(fn));
    if (skip > size)
        fail(GCT_ERR_RUNTIME, "Error during the reading of file " + fn + ".\nThe actual file size :: " +
                                  std::to_string(size) + " bytes.\nThe number of bytes to be skipped :: " +
                                  std::to_string(skip) + " bytes.");
    if (bytes != size - skip)
        fail(GCT_ERR_RUNTIME, "Error during the reading of file " + fn + ".\nThe actual file size :: " +
                                  std::to_string(size) + " bytes.\nThe number of bytes requested :: " +
                                  std::to_string(bytes) + " bytes.");
    std::vector<uint8_t> out(bytes);
    f.seekg(std::streamoff(skip));
    f.read(reinterpret_cast<char *>(out.data()), std::streamsize(bytes));
    return out;
}

// .mha data: everything after the "ElementDataFile = LOCAL" line; like the
// reference, a short file leaves the missing elements at zero
// (meta_img_reader.cpp:145-163).
std::vector<uint8_t> read_mha_local(const std::string &fn, uint64_t bytes)
{
    std::ifstream f(fn, std::ios::binary);
    if (!f) fail(GCT_ERR_RUNTIME, "File " + fn + " cannot be opened for reading data.");
    std::string line;
    bool found = false;
    while (std::getline(f, line)) {
        if (line == "ElementDataFile = LOCAL" || line == "ElementDataFile = LOCAL\r") { found = true; break; }
    }
    if (!found) fail(GCT_ERR_RUNTIME, "File " + fn + " has no 'ElementDataFile = LOCAL' line.");
    std::vector<uint8_t> out(bytes, 0);
    f.read(reinterpret_cast<char *>(out.data()), std::streamsize(bytes));
    return out;
}

// Every file under the directory of `fn` (recursively) with its extension and
// its size, in name order (file_utils.cpp:7-28).
std::vector<std::string> series_of(const std::string &fn)
{
    require_file(fn);
    const fs::path p(fn);
    const auto ext = p.extension();
    const auto size = fs::file_size(p);
    fs::path dir = p.parent_path();
    if (dir.empty()) dir = ".";
    std::vector<std::string> files;
    for (const auto &e : fs::recursive_directory_iterator(dir))
        if (e.is_regular_file() && e.path().extension() == ext && fs::file_size(e.path()) == size)
            files.push_back(e.path().string());
    std::sort(files.begin(), files.end());
    return files;
}

template <class I, class O>
void convert_as(const uint8_t *src, uint64_t n, O *dst)
{
    for (uint64_t i = 0; i < n; ++i) {
        I v;
        std::memcpy(&v, src + i * sizeof(I), sizeof(I));
        dst[i] = static_cast<O>(v);
    }
}

template <class O>
void convert_to(const uint8_t *src, int type, uint64_t n, O *dst)
{
    switch (type) {
    case GCT_T_UINT8: convert_as<uint8_t>(src, n, dst); break;
    case GCT_T_INT8: convert_as<int8_t>(src, n, dst); break;
    case GCT_T_UINT16: convert_as<uint16_t>(src, n, dst); break;
    case GCT_T_INT16: convert_as<int16_t>(src, n, dst); break;
    case GCT_T_UINT32: convert_as<uint32_t>(src, n, dst); break;
    case GCT_T_INT32: convert_as<int32_t>(src, n, dst); break;
    case GCT_T_UINT64: convert_as<uint64_t>(src, n, dst); break;
    case GCT_T_INT64: convert_as<int64_t>(src, n, dst); break;
    case GCT_T_FLOAT: convert_as<float>(src, n, dst); break;
    default: convert_as<double>(src, n, dst); break;
    }
}

void convert(const uint8_t *src, int in_type, uint64_t n, int out_type, void *dst)
{
    switch (out_type) {
    case GCT_T_UINT8: convert_to(src, in_type, n, static_cast<uint8_t *>(dst)); break;
    case GCT_T_INT8: convert_to(src, in_type, n, static_cast<int8_t *>(dst)); break;
    case GCT_T_UINT16: convert_to(src, in_type, n, static_cast<uint16_t *>(dst)); break;
    case GCT_T_INT16: convert_to(src, in_type, n, static_cast<int16_t *>(dst)); break;
    case GCT_T_UINT32: convert_to(src, in_type, n, static_cast<uint32_t *>(dst)); break;
    case GCT_T_INT32: convert_to(src, in_type, n, static_cast<int32_t *>(dst)); break;
    case GCT_T_UINT64: convert_to(src, in_type, n, static_cast<uint64_t *>(dst)); break;
    case GCT_T_INT64: convert_to(src, in_type, n, static_cast<int64_t *>(dst)); break;
    case GCT_T_FLOAT: convert_to(src, in_type, n, static_cast<float *>(dst)); break;
    case GCT_T_DOUBLE: convert_to(src, in_type, n, static_cast<double *>(dst)); break;
    default: fail(GCT_ERR_INVALID_ARGUMENT, "unknown element type");
    }
}

// Directory of `filename` must exist and be writable by its owner
// (file_utils.h:20-30).  A bare file name means the working directory (the
// reference rejects it: its parent path is empty).
void require_writable_parent(const std::string &filename, const char *what)
{
    fs::path dir = fs::path(filename).parent_path();
    if (dir.empty()) dir = ".";
    const auto st = fs::status(dir);
    if (!fs::exists(st) || !fs::is_directory(st) || (st.permissions() & fs::perms::owner_write) == fs::perms::none)
        fail(GCT_ERR_RUNTIME, std::string("Error during the writing of ") + what + " " + filename +
                                  "\nParent directory is not writable.");
}

void write_bytes(const std::string &filename, const void *data, uint64_t bytes, const char *what)
{
    std::ofstream f(filename, std::ios::binary | std::ios::app);
    if (!f) fail(GCT_ERR_RUNTIME, std::string("Error during the writing of ") + what + " " + filename +
                                      "\nFile cannot be opened.");
    f.write(static_cast<const char *>(data), std::streamsize(bytes));
    if (!f) fail(GCT_ERR_RUNTIME, std::string("Error during the writing of ") + what + " " + filename);
}

void truncate_new(const std::string &filename, const char *what)
{
    std::ofstream f(filename, std::ios::binary | std::ios::trunc);
    if (!f) fail(GCT_ERR_RUNTIME, std::string("Error during the writing of ") + what + " " + filename +
                                      "\nFile cannot be opened.");
}

// std::to_string of a float: "%f" (meta_img_file_writer.h:61-71).
std::string join_floats(const float *v, uint32_t n)
{
    std::string s;
    for (uint32_t i = 0; i < n; ++i) {
        if (i) s += " ";
        s += std::to_string(v[i]);
    }
    return s;
}

}  // namespace
}  // namespace gct

using namespace gct;

extern "C" gct_status gct_file_read_metadata(const char *filename, int32_t format, const gct_array_metadata *raw_md,
                                             gct_array_metadata *md)
{
    return guarded([&] {
        if (!filename || !md) fail(GCT_ERR_INVALID_ARGUMENT, "null argument");
        *md = read_md(filename, format_of(filename, format), raw_md);
    });
}

extern "C" gct_status gct_file_read(const char *filename, int32_t format, int series, const gct_array_metadata *raw_md,
                                    int32_t out_type, void *out, uint64_t capacity, uint64_t *count)
{
    return guarded([&] {
        if (!filename || !count) fail(GCT_ERR_INVALID_ARGUMENT, "null argument");
        (void)elem_size(out_type);
        const std::string fn(filename);
        const int fmt = format_of(fn, format);
        const gct_array_metadata md = read_md(fn, fmt, raw_md);
        const uint64_t n1 = md_elems(md);
        const uint64_t b1 = n1 * elem_size(md.type);
        // files holding the elements: the data file of a MetaImage header, the file itself otherwise
        const std::string first = (fmt == GCT_FILE_META) ? meta_data_file(fn) : fn;
        const std::vector<std::string> files = series ? series_of(first) : std::vector<std::string>{first};
        *count = n1 * files.size();
        if (!out) return;   // size query
        if (capacity < *count)
            fail(GCT_ERR_INVALID_ARGUMENT, "output holds " + std::to_string(capacity) + " elements, the file(s) " +
                                               std::to_string(*count));
        uint8_t *dst = static_cast<uint8_t *>(out);
        const size_t osz = elem_size(out_type);
        const bool local = fmt == GCT_FILE_META && first == fn;   // .mha: every file carries its header
        for (size_t k = 0; k < files.size(); ++k) {
            std::vector<uint8_t> bytes;
            if (local) bytes = read_mha_local(files[k], b1);
            else bytes = read_exact(files[k], b1, fmt == GCT_FILE_DEN ? 3 * sizeof(uint16_t) : 0);
            convert(bytes.data(), md.type, n1, out_type, dst + k * n1 * osz);
        }
    });
}

extern "C" const char *gct_meta_type_name(int32_t type)
{
    for (const MetaName &m : kMetaNames)
        if (m.type == type) return m.name;
    return nullptr;
}

extern "C" gct_status gct_meta_write(const char *filename, uint32_t ndims, const uint32_t *dim_size,
                                     const float *spacing, const float *offset, int32_t type, const void *data,
                                     uint64_t num_elems)
{
    return guarded([&] {
        if (!filename || !dim_size || !spacing || !offset || (!data && num_elems))
            fail(GCT_ERR_INVALID_ARGUMENT, "null argument");
        if (ndims < 1 || ndims > 3) fail(GCT_ERR_INVALID_ARGUMENT, "MetaImage files here hold 1 to 3 dimensions");
        const char *tname = gct_meta_type_name(type);
        if (!tname) fail(GCT_ERR_INVALID_ARGUMENT, "unknown element type");
        const std::string fn(filename);
        uint64_t n = 1;
        for (uint32_t i = 0; i < ndims; ++i) n *= dim_size[i];
        if (n != num_elems)
            fail(GCT_ERR_RUNTIME, "Error during the writing of Meta image file " + fn +
                                      "\nMismatch in the number of elements between data and metadata.");
        require_writable_parent(fn, "Meta image file");
        const fs::path path(fn);
        const bool local = lower(path.extension().string()) == ".mha";
        std::string dims;
        for (uint32_t i = 0; i < ndims; ++i) dims += (i ? " " : "") + std::to_string(dim_size[i]);
        const std::string data_name = local ? std::string("LOCAL") : path.stem().string() + ".raw";
        std::ostringstream h;
        h << "ObjectType = Image\n"
          << "NDims = " << ndims << "\n"
          << "BinaryData = True\n"
          << "DimSize = " << dims << "\n"
          << "ElementSpacing = " << join_floats(spacing, ndims) << "\n"
          << "Offset = " << join_floats(offset, ndims) << "\n"
          << "ElementType = " << tname << "\n"
          << "BinaryDataByteOrderMSB = False\n"
          << "ElementDataFile = " << data_name << "\n";
        const std::string hs = h.str();
        truncate_new(fn, "Meta image file");
        write_bytes(fn, hs.data(), hs.size(), "Meta image file");
        const uint64_t bytes = n * elem_size(type);
        if (local) {
            write_bytes(fn, data, bytes, "Meta image file");
        } else {
            const std::string raw = (path.parent_path() / data_name).string();
            require_writable_parent(raw, "binary file");
            truncate_new(raw, "binary file");
            write_bytes(raw, data, bytes, "binary file");
        }
    });
}

extern "C" gct_status gct_den_write(const char *filename, uint32_t num_rows, uint32_t num_cols, uint32_t num_slices,
                                    int32_t type, const void *data, uint64_t num_elems)
{
    return guarded([&] {
        if (!filename || (!data && num_elems)) fail(GCT_ERR_INVALID_ARGUMENT, "null argument");
        const std::string fn(filename);
        if (num_rows > 0xFFFFu || num_cols > 0xFFFFu || num_slices > 0xFFFFu)
            fail(GCT_ERR_INVALID_ARGUMENT, "Dennerlein sizes are 16-bit");
        const size_t es = elem_size(type);
        if (uint64_t(num_rows) * num_cols * num_slices != num_elems)
            fail(GCT_ERR_RUNTIME, "Error during the writing of Dennerlein file " + fn +
                                      "\nMismatch in the number of elements between data and metadata.");
        require_writable_parent(fn, "Dennerlein file");
        const uint16_t h[3] = {uint16_t(num_rows), uint16_t(num_cols), uint16_t(num_slices)};
        truncate_new(fn, "Dennerlein file");
        write_bytes(fn, h, sizeof(h), "Dennerlein file");
        write_bytes(fn, data, num_elems * es, "Dennerlein file");
    });
}

extern "C" gct_status gct_raw_write(const char *filename, int32_t type, const void *data, uint64_t num_elems)
{
    return guarded([&] {
        if (!filename || (!data && num_elems)) fail(GCT_ERR_INVALID_ARGUMENT, "null argument");
        const std::string fn(filename);
        const size_t es = elem_size(type);
        require_writable_parent(fn, "binary file");
        truncate_new(fn, "binary file");
        write_bytes(fn, data, num_elems * es, "binary file");
    });
}
