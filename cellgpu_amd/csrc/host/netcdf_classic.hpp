// Minimal NetCDF classic-format file (CDF-1 written; CDF-1 and CDF-2 read), the
// on-disk format of the reference's NcFile databases (netCDF legacy C++ API,
// classic mode). Only what those databases use: fixed dimensions plus one
// unlimited record dimension, record variables of type int or double, no
// attributes written (attributes of read files are skipped).
//
// Layout (Unidata "NetCDF Classic Format Specification"): a header (magic, record
// count, dimension list, global attributes, variable list with per-variable
// record slab size and offset), then the records, each the concatenation of
// every record variable's slab; all numbers big-endian.
#pragma once

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

namespace cgpu_host {

class NcClassic {
   public:
    enum Type { NC_INT = 4, NC_FLOAT = 5, NC_DOUBLE = 6 };  // classic-format type codes
    struct Dim {
        std::string name;
        int64_t len;  // 0: the record dimension
    };
    struct Var {
        std::string name;
        std::vector<int> dims;  // dimension ids; dims[0] the record dimension
        Type type;
        int64_t vsize;   // bytes of one record's slab
        int64_t begin;   // offset of record 0's slab
        int64_t count;   // values per record
    };

    // create (replacing) a file with these dimensions and record variables
    static NcClassic create(const std::string &path, const std::vector<Dim> &dims,
                            const std::vector<std::pair<std::string, std::pair<Type, std::vector<int>>>> &vars);
    // open an existing file; writable: records may be appended or overwritten
    static NcClassic open(const std::string &path, bool writable);

    NcClassic() = default;
    NcClassic(NcClassic &&o) noexcept;
    NcClassic &operator=(NcClassic &&o) noexcept;
    NcClassic(const NcClassic &) = delete;
    ~NcClassic();

    int64_t num_records() const { return numrecs_; }
    const std::vector<Dim> &dims() const { return dims_; }
    const std::vector<Var> &vars() const { return vars_; }
    int var_id(const std::string &name) const;  // -1 when absent
    int64_t dim_len(const std::string &name) const;  // -1 when absent

    // record rec (== num_records() appends) of variable v from doubles / ints;
    // values are converted to the variable's type
    void put_record(int v, int64_t rec, const double *vals);
    void put_record(int v, int64_t rec, const int *vals);
    void get_record(int v, int64_t rec, double *vals) const;
    void get_record(int v, int64_t rec, int *vals) const;
    void sync();

   private:
    void write_bytes(int64_t off, const std::vector<unsigned char> &b);
    std::vector<unsigned char> read_bytes(int64_t off, int64_t n) const;
    void set_numrecs(int64_t n);
    std::FILE *f_ = nullptr;
    bool writable_ = false;
    int version_ = 1;
    int64_t numrecs_ = 0, recsize_ = 0;
    std::vector<Dim> dims_;
    std::vector<Var> vars_;
};

}  // namespace cgpu_host
