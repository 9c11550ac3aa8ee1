// perphase_main.cpp -- a C++ host of libaxs3d.so with no Python: reads a
// small domain (raw little-endian arrays named in domain.txt, written by
// tests/test_capi_cpp.py), makes the set-up calls of include/axs3d.h
// itself, runs the per-phase Newmark loop through GpuDomain and writes the
// seismograms (FP64, [sample][receiver][3]) to a file.
//
// usage: axs3d_perphase <domain dir> <fp64|fp32> <nsteps> <out file>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "gpu_domain.hpp"

namespace {

struct DomainFiles {
    std::string dir;
    std::map<std::string, double> scalar;
    std::map<std::string, std::vector<char>> array;

    explicit DomainFiles(const std::string &d) : dir(d) {
        std::ifstream man(dir + "/domain.txt");
        if (!man) throw std::runtime_error("no " + dir + "/domain.txt");
        std::string line;
        while (std::getline(man, line)) {
            std::istringstream ls(line);
            std::string kind, name;
            ls >> kind >> name;
            if (kind == "scalar") {
                double v;
                ls >> v;
                scalar[name] = v;
            } else if (kind == "array") {
                std::size_t nbytes;
                ls >> nbytes;
                std::ifstream f(dir + "/" + name + ".bin", std::ios::binary);
                std::vector<char> buf(nbytes);
                if (!f.read(buf.data(), (std::streamsize)nbytes))
                    throw std::runtime_error("short array " + name);
                array[name] = std::move(buf);
            }
        }
    }
    bool has(const std::string &n) const { return array.count(n) != 0; }
    template <typename T>
    const T *ptr(const std::string &n) const {
        auto it = array.find(n);
        if (it == array.end()) throw std::runtime_error("missing " + n);
        return reinterpret_cast<const T *>(it->second.data());
    }
    template <typename T>
    int count(const std::string &n) const {
        return (int)(array.at(n).size() / sizeof(T));
    }
    int geti(const std::string &n) const { return (int)scalar.at(n); }
};

// the set-up calls, in the order the library requires (points, elements,
// materials, axis, source, receivers, dt, finalize)
void release(const GpuDomain &dom, const DomainFiles &f, int max_samples) {
    axs3d_ctx *c = dom.context();
    dom.check(axs3d_set_gmat(c, f.ptr<double>("g_gll"),
                             f.ptr<double>("g_glj")), "set_gmat");
    dom.check(axs3d_set_solid_points(c, f.count<int32_t>("solid_nr"),
                                     f.ptr<int32_t>("solid_nr"),
                                     f.ptr<double>("solid_mass")),
              "set_solid_points");
    const int ne = f.count<uint8_t>("elem_axial");
    dom.check(axs3d_set_solid_elements(c, ne, f.ptr<int32_t>("elem_points"),
                                       f.ptr<uint8_t>("elem_axial"),
                                       f.ptr<double>("elem_gq")),
              "set_solid_elements");
    dom.check(axs3d_set_solid_iso1d(c, f.ptr<double>("elem_iso")),
              "set_solid_iso1d");
    if (f.has("elem_kind")) {
        dom.check(axs3d_set_element_kinds(c, f.ptr<uint8_t>("elem_kind")),
                  "set_element_kinds");
        dom.check(axs3d_set_solid_iso3d(c, 0, ne, f.ptr<double>("prop3d")),
                  "set_solid_iso3d");
    }
    dom.check(axs3d_set_axial_points(c, f.count<int32_t>("axial_points"),
                                     f.ptr<int32_t>("axial_points")),
              "set_axial_points");
    dom.check(axs3d_set_moment_source(c, f.geti("src_elem"),
                                      f.geti("src_nu1"),
                                      f.ptr<double>("src_pattern"),
                                      f.count<double>("stf"),
                                      f.ptr<double>("stf")),
              "set_moment_source");
    dom.check(axs3d_set_receivers(c, f.count<int32_t>("rec_elem"),
                                  f.ptr<int32_t>("rec_elem"),
                                  f.ptr<double>("rec_weights"),
                                  f.ptr<double>("rec_phi"), max_samples),
              "set_receivers");
    dom.check(axs3d_set_dt(c, f.scalar.at("dt")), "set_dt");
    dom.check(axs3d_finalize(c), "finalize");
}

}  // namespace

int main(int argc, char **argv) {
    if (argc != 5) {
        std::fprintf(stderr,
                     "usage: %s <domain dir> <fp64|fp32> <nsteps> <out>\n",
                     argv[0]);
        return 2;
    }
    const int prec = std::strcmp(argv[2], "fp32") == 0 ? AXS3D_FP32_FP64
                                                       : AXS3D_FP64;
    const int nsteps = std::atoi(argv[3]);
    try {
        const DomainFiles files(argv[1]);
        GpuDomain domain(0, prec);
        release(domain, files, nsteps);
        solveNewmark(domain, nsteps, 50);
        const int nrec = files.count<int32_t>("rec_elem");
        std::vector<double> seis((std::size_t)nsteps * nrec * 3);
        domain.check(axs3d_read_seismograms(domain.context(), seis.data(),
                                            nsteps),
                     "read_seismograms");
        std::FILE *out = std::fopen(argv[4], "wb");
        if (!out || std::fwrite(seis.data(), sizeof(double), seis.size(),
                                out) != seis.size())
            throw std::runtime_error(std::string("cannot write ") + argv[4]);
        std::fclose(out);
        std::printf("axs3d_perphase: %d steps, %d receivers, %s\n", nsteps,
                    nrec, argv[2]);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "axs3d_perphase: %s\n", e.what());
        return 1;
    }
    return 0;
}
