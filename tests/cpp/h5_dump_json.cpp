// h5_dump_json.cpp -- walks an HDF5 file and prints its groups, datasets and attributes as JSON
// (test infrastructure: the reader side of the openPMD output tests; independent of the writer).
//
//   h5_dump_json FILE OUTDIR
//
// stdout: {"<path>": {"kind": "group"|"dataset", "attrs": {"<name>": {"dtype": D, "value": V}},
//                     "dtype": D, "shape": [...], "data": "<file>"}, ...}
// D is a numpy-style code ("f4", "f8", "u4", "u8", "i4", ..., "str"); dataset contents are written
// to OUTDIR/<k>.bin in that dtype (native byte order).
#include <hdf5.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

std::string esc(const std::string& s)
{
    std::string o;
    for (char c : s) {
        if (c == '"' || c == '\\') o += '\\';
        o += c;
    }
    return o;
}

std::string dtype_code(hid_t t)
{
    const H5T_class_t cls = H5Tget_class(t);
    const size_t size = H5Tget_size(t);
    if (cls == H5T_STRING) return "str";
    if (cls == H5T_FLOAT) return "f" + std::to_string(size);
    if (cls == H5T_INTEGER) return std::string(H5Tget_sign(t) == H5T_SGN_NONE ? "u" : "i") + std::to_string(size);
    return "other";
}

struct Ctx {
    std::string outdir;
    std::vector<std::string> entries;
    int nfiles = 0;
};

std::string attr_json(hid_t a)
{
    const hid_t ft = H5Aget_type(a);
    const hid_t sp = H5Aget_space(a);
    const std::string code = dtype_code(ft);
    std::string val;
    if (code == "str") {
        if (H5Tis_variable_str(ft) > 0) {
            char* p = nullptr;
            const hid_t mt = H5Tcopy(H5T_C_S1);
            H5Tset_size(mt, H5T_VARIABLE);
            H5Aread(a, mt, &p);
            val = "\"" + esc(p ? p : "") + "\"";
            H5free_memory(p);
            H5Tclose(mt);
        } else {
            const size_t n = H5Tget_size(ft);
            std::vector<char> buf(n + 1, 0);
            const hid_t mt = H5Tcopy(H5T_C_S1);
            H5Tset_size(mt, n + 1);
            H5Aread(a, mt, buf.data());
            val = "\"" + esc(buf.data()) + "\"";
            H5Tclose(mt);
        }
    } else {
        const hssize_t n = H5Sget_simple_extent_npoints(sp);
        const int rank = H5Sget_simple_extent_ndims(sp);
        std::vector<double> v((size_t)(n > 0 ? n : 1));
        H5Aread(a, H5T_NATIVE_DOUBLE, v.data());
        char num[64];
        if (rank == 0) {
            std::snprintf(num, sizeof(num), "%.17g", v[0]);
            val = num;
        } else {
            val = "[";
            for (hssize_t i = 0; i < n; ++i) {
                std::snprintf(num, sizeof(num), "%s%.17g", i ? ", " : "", v[(size_t)i]);
                val += num;
            }
            val += "]";
        }
    }
    H5Sclose(sp);
    H5Tclose(ft);
    return "{\"dtype\": \"" + code + "\", \"value\": " + val + "}";
}

herr_t on_attr(hid_t loc, const char* name, const H5A_info_t*, void* data)
{
    auto* out = static_cast<std::vector<std::string>*>(data);
    const hid_t a = H5Aopen(loc, name, H5P_DEFAULT);
    out->push_back("\"" + esc(name) + "\": " + attr_json(a));
    H5Aclose(a);
    return 0;
}

std::string attrs_of(hid_t obj)
{
    std::vector<std::string> items;
    H5Aiterate2(obj, H5_INDEX_NAME, H5_ITER_INC, nullptr, on_attr, &items);
    std::string s = "{";
    for (size_t i = 0; i < items.size(); ++i) s += (i ? ", " : "") + items[i];
    return s + "}";
}

void describe(hid_t file, const std::string& path, Ctx& c)
{
    const hid_t obj = H5Oopen(file, path.c_str(), H5P_DEFAULT);
    if (obj < 0) return;
    const H5I_type_t kind = H5Iget_type(obj);
    std::string e = "\"" + esc(path) + "\": {\"kind\": ";
    if (kind == H5I_GROUP) {
        e += "\"group\", \"attrs\": " + attrs_of(obj) + "}";
    } else if (kind == H5I_DATASET) {
        const hid_t ft = H5Dget_type(obj);
        const hid_t mt = H5Tget_native_type(ft, H5T_DIR_ASCEND);
        const hid_t sp = H5Dget_space(obj);
        const int rank = H5Sget_simple_extent_ndims(sp);
        std::vector<hsize_t> dims((size_t)(rank > 0 ? rank : 1), 1);
        if (rank > 0) H5Sget_simple_extent_dims(sp, dims.data(), nullptr);
        const hssize_t n = H5Sget_simple_extent_npoints(sp);
        std::vector<char> buf((size_t)n * H5Tget_size(mt) + 1);
        if (n > 0) H5Dread(obj, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.data());
        const std::string fname = std::to_string(c.nfiles++) + ".bin";
        FILE* f = std::fopen((c.outdir + "/" + fname).c_str(), "wb");
        if (f) {
            std::fwrite(buf.data(), 1, (size_t)n * H5Tget_size(mt), f);
            std::fclose(f);
        }
        std::string shape = "[";
        for (int i = 0; i < rank; ++i) shape += (i ? ", " : "") + std::to_string((unsigned long long)dims[(size_t)i]);
        shape += "]";
        e += "\"dataset\", \"dtype\": \"" + dtype_code(ft) + "\", \"shape\": " + shape + ", \"data\": \"" + fname +
             "\", \"attrs\": " + attrs_of(obj) + "}";
        H5Sclose(sp);
        H5Tclose(mt);
        H5Tclose(ft);
    } else {
        e += "\"other\"}";
    }
    H5Oclose(obj);
    c.entries.push_back(e);
}

herr_t on_link(hid_t, const char* name, const H5L_info_t*, void* data)
{
    auto* c = static_cast<std::pair<hid_t, Ctx*>*>(data);
    describe(c->first, std::string("/") + name, *c->second);
    return 0;
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc != 3) {
        std::fprintf(stderr, "usage: h5_dump_json FILE OUTDIR\n");
        return 2;
    }
    const hid_t file = H5Fopen(argv[1], H5F_ACC_RDONLY, H5P_DEFAULT);
    if (file < 0) {
        std::fprintf(stderr, "cannot open %s\n", argv[1]);
        return 1;
    }
    Ctx c;
    c.outdir = argv[2];
    describe(file, "/", c);
    std::pair<hid_t, Ctx*> data(file, &c);
    H5Lvisit(file, H5_INDEX_NAME, H5_ITER_INC, on_link, &data);
    H5Fclose(file);
    std::printf("{\n");
    for (size_t i = 0; i < c.entries.size(); ++i) std::printf("%s%s\n", c.entries[i].c_str(), i + 1 < c.entries.size() ? "," : "");
    std::printf("}\n");
    return 0;
}
