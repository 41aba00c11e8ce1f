// h5_dump.cpp - prints every dataset of an HDF5 file as JSON (shape, element
// class and size, chunk shape, deflate level, values) for
// tests/test_trajectory_h5.py.  Test infrastructure only.
//
// usage: h5_dump FILE
#include <hdf5.h>

#include <cstdio>
#include <string>
#include <vector>

namespace {

void print_json_string(const char* s) {
    std::putchar('"');
    for (const char* p = s; p && *p; ++p) {
        if (*p == '"' || *p == '\\') std::putchar('\\');
        std::putchar(*p);
    }
    std::putchar('"');
}

bool first_ds = true;

herr_t visit(hid_t root, const char* name, const H5L_info_t*, void*) {
    const hid_t ds = H5Dopen2(root, name, H5P_DEFAULT);
    if (ds < 0) return 0;  // a group
    const hid_t space = H5Dget_space(ds);
    const int rank = H5Sget_simple_extent_ndims(space);
    std::vector<hsize_t> dims(rank > 0 ? rank : 1, 1);
    if (rank > 0) H5Sget_simple_extent_dims(space, dims.data(), nullptr);
    hsize_t count = 1;
    for (int r = 0; r < rank; ++r) count *= dims[r];
    const hid_t type = H5Dget_type(ds);
    const H5T_class_t cls = H5Tget_class(type);
    const size_t size = H5Tget_size(type);
    const bool varstr = cls == H5T_STRING && H5Tis_variable_str(type) > 0;
    const hid_t dcpl = H5Dget_create_plist(ds);
    std::printf("%s\"/%s\": {\"shape\": [", first_ds ? "" : ", ", name);
    first_ds = false;
    for (int r = 0; r < rank; ++r) std::printf("%s%llu", r ? ", " : "", static_cast<unsigned long long>(dims[r]));
    const char* cname = cls == H5T_FLOAT ? "float" : cls == H5T_INTEGER ? "int" : cls == H5T_STRING ? "string" : "other";
    int sign = cls == H5T_INTEGER ? (H5Tget_sign(type) == H5T_SGN_NONE ? 0 : 1) : -1;
    std::printf("], \"class\": \"%s\", \"size\": %zu, \"signed\": %d, \"chunk\": ", cname, varstr ? 0 : size, sign);
    if (H5Pget_layout(dcpl) == H5D_CHUNKED) {
        std::vector<hsize_t> ch(rank > 0 ? rank : 1);
        H5Pget_chunk(dcpl, rank, ch.data());
        std::printf("[");
        for (int r = 0; r < rank; ++r) std::printf("%s%llu", r ? ", " : "", static_cast<unsigned long long>(ch[r]));
        std::printf("]");
    } else {
        std::printf("null");
    }
    int deflate = -1;
    const int nf = H5Pget_nfilters(dcpl);
    for (int f = 0; f < nf; ++f) {
        unsigned flags = 0, cd[4] = {0, 0, 0, 0};
        size_t ncd = 4;
        char fname[64];
        const H5Z_filter_t id = H5Pget_filter2(dcpl, static_cast<unsigned>(f), &flags, &ncd, cd, sizeof(fname),
                                                fname, nullptr);
        if (id == H5Z_FILTER_DEFLATE) deflate = ncd ? static_cast<int>(cd[0]) : 0;
    }
    std::printf(", \"deflate\": %d, \"data\": [", deflate);
    if (cls == H5T_FLOAT) {
        std::vector<double> v(count);
        if (count) H5Dread(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data());
        for (hsize_t k = 0; k < count; ++k) std::printf("%s%.17g", k ? ", " : "", v[k]);
    } else if (cls == H5T_INTEGER) {
        std::vector<long long> v(count);
        if (count) H5Dread(ds, H5T_NATIVE_LLONG, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data());
        for (hsize_t k = 0; k < count; ++k) std::printf("%s%lld", k ? ", " : "", v[k]);
    } else if (varstr) {
        std::vector<char*> v(count, nullptr);
        if (count) H5Dread(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data());
        for (hsize_t k = 0; k < count; ++k) {
            if (k) std::printf(", ");
            print_json_string(v[k]);
        }
        if (count) H5Dvlen_reclaim(type, space, H5P_DEFAULT, v.data());
    }
    std::printf("]}");
    H5Pclose(dcpl);
    H5Tclose(type);
    H5Sclose(space);
    H5Dclose(ds);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 2) {
        std::fprintf(stderr, "usage: h5_dump FILE\n");
        return 2;
    }
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    const hid_t f = H5Fopen(argv[1], H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) {
        std::fprintf(stderr, "cannot open %s\n", argv[1]);
        return 1;
    }
    std::printf("{");
    H5Lvisit(f, H5_INDEX_NAME, H5_ITER_INC, visit, nullptr);
    std::printf("}\n");
    H5Fclose(f);
    return 0;
}
