// Replays a recorded single-particle proposal stream through the C++ EnergyTerm-shaped
// adapters (faunus_amd/host/gpu_energy.hpp) — the way the reference's
// MetropolisMonteCarlo::performMove would call GpuNonbonded/GpuEwald — and writes
// the per-move energy change and decision.
//
// usage: adapter_replay <dir>   with dir/config.json, dir/system.bin, dir/moves.bin;
// writes dir/du.bin (double) and dir/accepted.bin (uint8), prints the drift.
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../faunus_amd/host/gpu_energy.hpp"

template <class T> static std::vector<T> read_block(std::ifstream& f, size_t count) {
    std::vector<T> v(count);
    f.read(reinterpret_cast<char*>(v.data()), static_cast<std::streamsize>(sizeof(T) * count));
    if (!f) throw std::runtime_error("short read");
    return v;
}

int main(int argc, char** argv) {
    if (argc != 2) {
        std::fprintf(stderr, "usage: adapter_replay <dir>\n");
        return 2;
    }
    const std::string dir = argv[1];
    try {
        std::ifstream cj(dir + "/config.json");
        std::stringstream ss;
        ss << cj.rdbuf();
        faunus_amd::Metropolis mc(ss.str());
        std::ifstream sys(dir + "/system.bin", std::ios::binary);
        const int n = read_block<int>(sys, 1)[0];
        auto xyz = read_block<double>(sys, 3 * static_cast<size_t>(n));
        auto q = read_block<double>(sys, static_cast<size_t>(n));
        auto ids = read_block<int>(sys, static_cast<size_t>(n));
        const double e0 = mc.init(n, xyz.data(), q.data(), ids.data(), {0}, {n}, {0});
        std::ifstream mv(dir + "/moves.bin", std::ios::binary);
        const long count = read_block<long>(mv, 1)[0];
        auto index = read_block<int>(mv, static_cast<size_t>(count));
        auto pos = read_block<double>(mv, 3 * static_cast<size_t>(count));
        auto u = read_block<double>(mv, static_cast<size_t>(count));
        std::vector<double> du(static_cast<size_t>(count));
        std::vector<unsigned char> acc(static_cast<size_t>(count));
        for (long k = 0; k < count; ++k) {
            faunus_amd::Change c;
            faunus_amd::Change::GroupChange g;
            g.group_index = 0;
            g.internal = true;
            g.relative_atom_indices = {index[static_cast<size_t>(k)]};
            c.groups.push_back(g);
            std::vector<double> p(pos.begin() + 3 * k, pos.begin() + 3 * k + 3);
            acc[static_cast<size_t>(k)] = mc.perform(c, {index[static_cast<size_t>(k)]}, p, u[static_cast<size_t>(k)],
                                                     &du[static_cast<size_t>(k)]) ? 1 : 0;
        }
        std::ofstream o1(dir + "/du.bin", std::ios::binary);
        o1.write(reinterpret_cast<const char*>(du.data()), static_cast<std::streamsize>(sizeof(double) * du.size()));
        std::ofstream o2(dir + "/accepted.bin", std::ios::binary);
        o2.write(reinterpret_cast<const char*>(acc.data()), static_cast<std::streamsize>(acc.size()));
        std::printf("initial %.12g drift %.3e\n", e0, mc.relative_energy_drift());
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
