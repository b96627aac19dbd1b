// wf_stencil_generator.cpp -- drop-in for the SAT and 2D-SOR generators
// (SAT/generator.cpp:10-74, identical in 2D-SOR/nested_loop): `generator <e1>
// <e2> [seed]` writes ./Data/x_2_<e>_y_2_<f>.txt for (e1,e2), (e1-1,e2-1), ...
// while the first exponent is >= WF_GEN_MIN_EXP (13 in the reference).  Each
// file is "n1 n2" and then n2 + 3 rows of n1 + 3 values in [0, 10000); the
// smaller files hold the top-left corners of the same (2^e2 + 3) x (2^e1 + 3)
// table, as in the reference.  Without a seed the wall clock seeds the
// generator, as srand(time(NULL)) does; with one, the files are reproducible
// (splitmix64 stream).
#include <sys/stat.h>

#include <cstdint>
#include <cstdlib>
#include <ctime>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#ifndef WF_GEN_MIN_EXP
#define WF_GEN_MIN_EXP 13
#endif

using namespace std;

static uint64_t splitmix64(uint64_t &s)
{
    uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

int main(int argc, char *argv[])
{
    if (argc < 3 || argc > 4) {
        cout << "Incorrect parameters have been passed." << endl;
        exit(0);
    }
    int n1 = atoi(argv[1]);
    int n2 = atoi(argv[2]);
    uint64_t seed = (argc >= 4) ? strtoull(argv[3], nullptr, 10) : (uint64_t)time(NULL);
    if (n1 < 0 || n2 < 0 || n1 > 15 || n2 > 15) {
        cout << "Incorrect parameters have been passed." << endl;
        exit(0);
    }
    const int pad = 3;
    const int64_t size1 = 1LL << n1, size2 = 1LL << n2;
    const int64_t w = size1 + pad;
    vector<int> arr((size_t)(w * (size2 + pad)));
    for (auto &v : arr) v = (int)(splitmix64(seed) % 10000u);
    mkdir("./Data", 0755);
    while (n1 >= WF_GEN_MIN_EXP && n2 >= 0) {
        const int64_t s1 = 1LL << n1, s2 = 1LL << n2;
        const string path = "./Data/x_2_" + to_string(n1) + "_y_2_" + to_string(n2) + ".txt";
        ofstream f(path.c_str());
        f << s1 << " " << s2 << endl;
        for (int64_t j = 0; j < s2 + pad; j++) {
            for (int64_t i = 0; i < s1 + pad; i++) f << arr[(size_t)(j * w + i)] << " ";
            f << endl;
        }
        n1--;
        n2--;
    }
    return 0;
}
