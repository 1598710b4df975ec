// Command-line access to include/bfps_hip_h5.hpp for the CPU tests
// (tests/test_h5_cpp.py), which compare it with the reference's golden
// checkpoint file and with the two Python readers.
//
//   h5_check keys   FILE GROUP              names in GROUP, one a line
//   h5_check info   FILE DSET               kind size dim0 dim1 ...
//   h5_check read   FILE DSET OUT           values as float64 (a complex value is two)
//   h5_check write  FILE DSET KIND SIZE IN DIM...   append (KIND f|i|c; IN raw bytes)
//   h5_check group  FILE GROUP              create a group (and its parents)
//   h5_check fresh  FILE                    an empty file
#include <cstdio>
#include <fstream>
#include <iostream>
#include <iterator>
#include <string>
#include <vector>

#include "bfps_hip_h5.hpp"

using namespace bfps_hip;

int main(int argc, char** argv) {
    if (argc < 3) {
        std::cerr << "usage: h5_check keys|info|read|write|group|fresh FILE ...\n";
        return 2;
    }
    const std::string cmd = argv[1], file = argv[2];
    try {
        if (cmd == "keys" && argc == 4) {
            h5::File f(file, 'r');
            for (const std::string& k : f.keys(argv[3])) std::cout << k << "\n";
        } else if (cmd == "info" && argc == 4) {
            h5::File f(file, 'r');
            const h5::dtype t = f.type(argv[3]);
            std::cout << t.kind << " " << t.size;
            for (uint64_t s : f.shape(argv[3])) std::cout << " " << s;
            std::cout << "\n";
        } else if (cmd == "read" && argc == 5) {
            h5::File f(file, 'r');
            const std::vector<double> v = f.read<double>(argv[3]);
            std::ofstream o(argv[4], std::ios::binary);
            o.write(reinterpret_cast<const char*>(v.data()), std::streamsize(v.size() * sizeof(double)));
        } else if (cmd == "write" && argc >= 7) {
            h5::dtype t{argv[4][0], size_t(std::stoul(argv[5]))};
            std::ifstream in(argv[6], std::ios::binary);
            std::vector<char> raw((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
            std::vector<uint64_t> shape;
            for (int i = 7; i < argc; i++) shape.push_back(std::stoull(argv[i]));
            h5::File f(file, 'a');
            f.write(argv[3], t, shape, raw.data());
            f.close();
        } else if (cmd == "group" && argc == 4) {
            h5::File f(file, 'a');
            f.create_group(argv[3]);
            f.close();
        } else if (cmd == "fresh" && argc == 3) {
            h5::File f(file, 'w');
            f.close();
        } else {
            std::cerr << "bad arguments\n";
            return 2;
        }
    } catch (const h5::error& e) {
        std::cerr << "h5 error: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
