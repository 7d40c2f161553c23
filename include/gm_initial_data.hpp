// gm_initial_data.hpp - load_initial_modes_binary (include/InitialData.hpp:18-34,
// src/InitialData.cpp of rdellamonica/boson-cloud-binary) for the MI355X
// solver's drop-in class: the MODESBIN initial state.
//
// Layout (little-endian): the 8 bytes "MODESBIN", int32 version (= 1),
// int64 Nr, int64 number of stored modes, then per mode int32 l, int32 m,
// Nr doubles of Re U_lm(r_i) and Nr doubles of Im U_lm(r_i). Stored modes
// the solver does not list are ignored (a later duplicate replaces an
// earlier one); listed modes that are not stored are zero.
#pragma once

#include <complex>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <map>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "gm_cnm.hpp"

namespace gm {

inline std::vector<std::complex<double>> load_initial_modes_binary(const std::string &filename,
                                                                    const std::vector<mode> &modes)
{
    std::ifstream f(filename, std::ios::binary);
    if (!f) throw std::runtime_error("Cannot open initial-state file " + filename);
    char magic[8];
    int32_t version = 0;
    int64_t Nr = 0, count = 0;
    if (!f.read(magic, 8) || std::memcmp(magic, "MODESBIN", 8) != 0)
        throw std::runtime_error(filename + " is not a MODESBIN file");
    if (!f.read(reinterpret_cast<char *>(&version), 4) || version != 1)
        throw std::runtime_error("Unsupported MODESBIN version in " + filename);
    if (!f.read(reinterpret_cast<char *>(&Nr), 8) || !f.read(reinterpret_cast<char *>(&count), 8) || Nr < 0 ||
        count < 0)
        throw std::runtime_error(filename + " is truncated");
    std::map<std::pair<int, int>, std::vector<std::complex<double>>> stored;
    std::vector<double> re(static_cast<size_t>(Nr)), im(static_cast<size_t>(Nr));
    for (int64_t k = 0; k < count; ++k) {
        int32_t lm[2];
        if (!f.read(reinterpret_cast<char *>(lm), 8) ||
            !f.read(reinterpret_cast<char *>(re.data()), static_cast<std::streamsize>(8 * Nr)) ||
            !f.read(reinterpret_cast<char *>(im.data()), static_cast<std::streamsize>(8 * Nr)))
            throw std::runtime_error(filename + " is truncated");
        auto &u = stored[{lm[0], lm[1]}];
        u.resize(static_cast<size_t>(Nr));
        for (int64_t i = 0; i < Nr; ++i) u[i] = {re[i], im[i]};
    }
    std::vector<std::complex<double>> psi(static_cast<size_t>(Nr) * modes.size());
    for (size_t j = 0; j < modes.size(); ++j) {
        auto it = stored.find({modes[j].l, modes[j].m});
        if (it != stored.end()) std::copy(it->second.begin(), it->second.end(), psi.begin() + j * Nr);
    }
    return psi;
}

}  // namespace gm
