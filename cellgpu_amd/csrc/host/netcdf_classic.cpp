// NetCDF classic-format encoder / decoder (see netcdf_classic.hpp).
#include "netcdf_classic.hpp"

#include <algorithm>
#include <climits>
#include <cstring>
#include <stdexcept>

namespace cgpu_host {

namespace {

constexpr uint32_t NC_DIMENSION = 0x0A, NC_VARIABLE = 0x0B, NC_ATTRIBUTE = 0x0C;
constexpr uint32_t STREAMING = 0xFFFFFFFFu;

int64_t pad4(int64_t n) { return (n + 3) & ~int64_t(3); }
int type_size(int t) {
    switch (t) {
        case 1: case 2: return 1;  // byte, char
        case 3: return 2;          // short
        case 4: case 5: return 4;  // int, float
        case 6: return 8;          // double
        default: throw std::runtime_error("netcdf: unknown type");
    }
}

struct Out {
    std::vector<unsigned char> b;
    void u32(uint32_t v) {
        for (int k = 3; k >= 0; --k) b.push_back((unsigned char)(v >> (8 * k)));
    }
    void u64(uint64_t v) {
        for (int k = 7; k >= 0; --k) b.push_back((unsigned char)(v >> (8 * k)));
    }
    void name(const std::string &s) {
        u32((uint32_t)s.size());
        for (char c : s) b.push_back((unsigned char)c);
        while (b.size() % 4) b.push_back(0);
    }
};

struct In {
    const std::vector<unsigned char> &b;
    size_t p = 0;
    void need(size_t n) const {
        if (p + n > b.size()) throw std::runtime_error("netcdf: truncated header");
    }
    uint32_t u32() {
        need(4);
        uint32_t v = 0;
        for (int k = 0; k < 4; ++k) v = (v << 8) | b[p++];
        return v;
    }
    uint64_t u64() {
        need(8);
        uint64_t v = 0;
        for (int k = 0; k < 8; ++k) v = (v << 8) | b[p++];
        return v;
    }
    std::string name() {
        uint32_t n = u32();
        need(pad4(n));
        std::string s(b.begin() + p, b.begin() + p + n);
        p += pad4(n);
        return s;
    }
    void skip_attrs() {
        uint32_t tag = u32(), n = u32();
        if (tag == 0 && n == 0) return;
        if (tag != NC_ATTRIBUTE) throw std::runtime_error("netcdf: bad attribute list");
        for (uint32_t a = 0; a < n; ++a) {
            name();
            int t = (int)u32();
            uint32_t m = u32();
            int64_t sz = pad4((int64_t)m * type_size(t));
            need(sz);
            p += sz;
        }
    }
};

void put_be(unsigned char *d, const void *src, int size) {
    const unsigned char *s = (const unsigned char *)src;
    for (int k = 0; k < size; ++k) d[k] = s[size - 1 - k];  // host is little-endian (x86-64)
}
void get_be(void *dst, const unsigned char *s, int size) {
    unsigned char *d = (unsigned char *)dst;
    for (int k = 0; k < size; ++k) d[k] = s[size - 1 - k];
}

}  // namespace

NcClassic::NcClassic(NcClassic &&o) noexcept { *this = std::move(o); }
NcClassic &NcClassic::operator=(NcClassic &&o) noexcept {
    if (this != &o) {
        if (f_) std::fclose(f_);
        f_ = o.f_;
        o.f_ = nullptr;
        writable_ = o.writable_;
        version_ = o.version_;
        numrecs_ = o.numrecs_;
        recsize_ = o.recsize_;
        dims_ = std::move(o.dims_);
        vars_ = std::move(o.vars_);
    }
    return *this;
}
NcClassic::~NcClassic() {
    if (f_) std::fclose(f_);
}

NcClassic NcClassic::create(const std::string &path, const std::vector<Dim> &dims,
                            const std::vector<std::pair<std::string, std::pair<Type, std::vector<int>>>> &vars) {
    NcClassic nc;
    nc.dims_ = dims;
    nc.version_ = 1;
    for (const auto &v : vars) {
        Var x;
        x.name = v.first;
        x.type = v.second.first;
        x.dims = v.second.second;
        if (x.dims.empty() || dims.at(x.dims[0]).len != 0) throw std::invalid_argument("netcdf: record variables only");
        x.count = 1;
        for (size_t k = 1; k < x.dims.size(); ++k) x.count *= dims.at(x.dims[k]).len;
        x.vsize = pad4(x.count * type_size(x.type));
        nc.vars_.push_back(x);
    }
    // header size: fixed part, names, then the variable entries
    Out h;
    h.b = {'C', 'D', 'F', 1};
    h.u32(0);
    h.u32(NC_DIMENSION);
    h.u32((uint32_t)dims.size());
    for (const auto &d : dims) {
        h.name(d.name);
        h.u32((uint32_t)d.len);
    }
    h.u32(0);  // no global attributes
    h.u32(0);
    h.u32(NC_VARIABLE);
    h.u32((uint32_t)nc.vars_.size());
    // first pass for the size, second with the offsets
    std::vector<size_t> begin_at;
    for (const auto &x : nc.vars_) {
        h.name(x.name);
        h.u32((uint32_t)x.dims.size());
        for (int d : x.dims) h.u32((uint32_t)d);
        h.u32(0);
        h.u32(0);
        h.u32((uint32_t)x.type);
        h.u32((uint32_t)x.vsize);
        begin_at.push_back(h.b.size());
        h.u32(0);
    }
    int64_t off = (int64_t)h.b.size();
    nc.recsize_ = 0;
    for (size_t k = 0; k < nc.vars_.size(); ++k) {
        nc.vars_[k].begin = off + nc.recsize_;
        nc.recsize_ += nc.vars_[k].vsize;
        if (nc.vars_[k].begin > 0x7FFFFFFF) throw std::runtime_error("netcdf: header beyond the CDF-1 offset range");
        uint32_t b = (uint32_t)nc.vars_[k].begin;
        for (int q = 0; q < 4; ++q) h.b[begin_at[k] + q] = (unsigned char)(b >> (8 * (3 - q)));
    }
    nc.f_ = std::fopen(path.c_str(), "w+b");
    if (!nc.f_) throw std::runtime_error("netcdf: cannot create " + path);
    nc.writable_ = true;
    nc.write_bytes(0, h.b);
    nc.sync();
    return nc;
}

NcClassic NcClassic::open(const std::string &path, bool writable) {
    NcClassic nc;
    nc.f_ = std::fopen(path.c_str(), writable ? "r+b" : "rb");
    if (!nc.f_) throw std::runtime_error("netcdf: cannot open " + path);
    nc.writable_ = writable;
    std::fseek(nc.f_, 0, SEEK_END);
    const int64_t fsize = std::ftell(nc.f_);
    // the header is read whole: it is at most the first megabytes of the file
    std::vector<unsigned char> b = nc.read_bytes(0, std::min<int64_t>(fsize, 1 << 24));
    In in{b};
    in.need(4);
    if (b[0] != 'C' || b[1] != 'D' || b[2] != 'F' || (b[3] != 1 && b[3] != 2))
        throw std::runtime_error("netcdf: not a classic-format file: " + path);
    nc.version_ = b[3];
    in.p = 4;
    uint32_t nrec = in.u32();
    uint32_t tag = in.u32(), n = in.u32();
    if (!(tag == 0 && n == 0)) {
        if (tag != NC_DIMENSION) throw std::runtime_error("netcdf: bad dimension list");
        for (uint32_t k = 0; k < n; ++k) {
            Dim d;
            d.name = in.name();
            d.len = in.u32();
            nc.dims_.push_back(d);
        }
    }
    in.skip_attrs();
    tag = in.u32();
    n = in.u32();
    if (!(tag == 0 && n == 0)) {
        if (tag != NC_VARIABLE) throw std::runtime_error("netcdf: bad variable list");
        for (uint32_t k = 0; k < n; ++k) {
            Var x;
            x.name = in.name();
            uint32_t nd = in.u32();
            for (uint32_t q = 0; q < nd; ++q) {
                uint32_t id = in.u32();
                if (id >= nc.dims_.size()) throw std::runtime_error("netcdf: bad dimension id");
                x.dims.push_back((int)id);
            }
            in.skip_attrs();
            x.type = (Type)in.u32();
            x.vsize = in.u32();
            x.begin = nc.version_ == 1 ? (int64_t)in.u32() : (int64_t)in.u64();
            x.count = 1;
            for (size_t q = 0; q < x.dims.size(); ++q)
                if (!(q == 0 && nc.dims_[x.dims[q]].len == 0)) x.count *= nc.dims_[x.dims[q]].len;
            type_size(x.type);
            nc.vars_.push_back(x);
        }
    }
    nc.recsize_ = 0;
    int nrecvars = 0;
    for (const auto &x : nc.vars_)
        if (!x.dims.empty() && nc.dims_[x.dims[0]].len == 0) {
            nc.recsize_ += x.vsize;
            ++nrecvars;
        }
    if (nrecvars == 1)  // a lone record variable's records are not padded
        for (const auto &x : nc.vars_)
            if (!x.dims.empty() && nc.dims_[x.dims[0]].len == 0) nc.recsize_ = x.count * type_size(x.type);
    if (nrec == STREAMING) {
        int64_t first = INT64_MAX;
        for (const auto &x : nc.vars_)
            if (!x.dims.empty() && nc.dims_[x.dims[0]].len == 0) first = std::min(first, x.begin);
        nc.numrecs_ = nc.recsize_ > 0 && first != INT64_MAX ? (fsize - first) / nc.recsize_ : 0;
    } else {
        nc.numrecs_ = nrec;
    }
    return nc;
}

int NcClassic::var_id(const std::string &name) const {
    for (size_t k = 0; k < vars_.size(); ++k)
        if (vars_[k].name == name) return (int)k;
    return -1;
}
int64_t NcClassic::dim_len(const std::string &name) const {
    for (const auto &d : dims_)
        if (d.name == name) return d.len;
    return -1;
}

void NcClassic::write_bytes(int64_t off, const std::vector<unsigned char> &b) {
    if (std::fseek(f_, (long)off, SEEK_SET) != 0 || std::fwrite(b.data(), 1, b.size(), f_) != b.size())
        throw std::runtime_error("netcdf: write failed");
}
std::vector<unsigned char> NcClassic::read_bytes(int64_t off, int64_t n) const {
    std::vector<unsigned char> b((size_t)n);
    if (std::fseek(f_, (long)off, SEEK_SET) != 0 || std::fread(b.data(), 1, b.size(), f_) != b.size())
        throw std::runtime_error("netcdf: read past the end of the file");
    return b;
}
void NcClassic::set_numrecs(int64_t n) {
    numrecs_ = n;
    Out o;
    o.u32((uint32_t)n);
    write_bytes(4, o.b);
}
void NcClassic::sync() {
    if (f_) std::fflush(f_);
}

void NcClassic::put_record(int v, int64_t rec, const double *vals) {
    if (!writable_) throw std::runtime_error("netcdf: file opened read-only");
    if (v < 0 || v >= (int)vars_.size()) throw std::invalid_argument("netcdf: no such variable");
    if (rec < 0 || rec > numrecs_) throw std::invalid_argument("netcdf: records are written in order");
    const Var &x = vars_[v];
    std::vector<unsigned char> b((size_t)x.vsize, 0);
    const int ts = type_size(x.type);
    for (int64_t k = 0; k < x.count; ++k) {
        if (x.type == NC_DOUBLE) {
            put_be(&b[k * ts], &vals[k], 8);
        } else if (x.type == NC_INT) {
            int32_t i = (int32_t)vals[k];
            put_be(&b[k * ts], &i, 4);
        } else {
            throw std::runtime_error("netcdf: only int and double variables are written");
        }
    }
    write_bytes(x.begin + rec * recsize_, b);
    if (rec == numrecs_) set_numrecs(numrecs_ + 1);
}
void NcClassic::put_record(int v, int64_t rec, const int *vals) {
    if (v < 0 || v >= (int)vars_.size()) throw std::invalid_argument("netcdf: no such variable");
    std::vector<double> d(vals, vals + vars_[v].count);
    put_record(v, rec, d.data());
}

void NcClassic::get_record(int v, int64_t rec, double *vals) const {
    if (v < 0 || v >= (int)vars_.size()) throw std::invalid_argument("netcdf: no such variable");
    if (rec < 0 || rec >= numrecs_) throw std::invalid_argument("netcdf: no such record");
    const Var &x = vars_[v];
    const int ts = type_size(x.type);
    std::vector<unsigned char> b = read_bytes(x.begin + rec * recsize_, x.count * ts);
    for (int64_t k = 0; k < x.count; ++k) {
        switch (x.type) {
            case NC_DOUBLE: get_be(&vals[k], &b[k * ts], 8); break;
            case NC_INT: {
                int32_t i;
                get_be(&i, &b[k * ts], 4);
                vals[k] = i;
                break;
            }
            case NC_FLOAT: {
                float f;
                get_be(&f, &b[k * ts], 4);
                vals[k] = f;
                break;
            }
            default: throw std::runtime_error("netcdf: unsupported variable type");
        }
    }
}
void NcClassic::get_record(int v, int64_t rec, int *vals) const {
    if (v < 0 || v >= (int)vars_.size()) throw std::invalid_argument("netcdf: no such variable");
    std::vector<double> d((size_t)vars_[v].count);
    get_record(v, rec, d.data());
    for (size_t k = 0; k < d.size(); ++k) vals[k] = (int)d[k];
}

}  // namespace cgpu_host
