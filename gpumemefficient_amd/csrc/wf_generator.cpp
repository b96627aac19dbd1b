// wf_generator.cpp -- drop-in for the reference's generator
// (LCS/generator.cpp:10-77): `generator <e1> <e2> [seed] [alphabet]` writes
// ./Data/str1_2_<e>_str2_2_<f>.txt for (e1,e2), (e1-1,e2-1), ... while the
// first exponent is >= WF_GEN_MIN_EXP (4 in LCS/SmithWaterman, 8 in
// EditDistance/LCS-SHFL).  Each smaller file holds the prefixes of the same
// two sequences, as in the reference.  Symbols are uniform in [0, alphabet)
// (alphabet 4 by default, like rand()%4).  Without a seed the wall clock
// seeds the generator, as srand(time(NULL)) does; with one, the files are
// reproducible (splitmix64 stream).
#include <sys/stat.h>

#include <cstdint>
#include <cstdlib>
#include <ctime>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#ifndef WF_GEN_MIN_EXP
#define WF_GEN_MIN_EXP 4
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
    if (argc < 3 || argc > 5) {
        cout << "Incorrect parameters have been passed." << endl;
        exit(0);
    }
    int n1 = atoi(argv[1]);
    int n2 = atoi(argv[2]);
    uint64_t seed = (argc >= 4) ? strtoull(argv[3], nullptr, 10) : (uint64_t)time(NULL);
    const int alphabet = (argc >= 5) ? atoi(argv[4]) : 4;
    if (n1 < 0 || n2 < 0 || n1 > 30 || n2 > 30 || alphabet < 1) {
        cout << "Incorrect parameters have been passed." << endl;
        exit(0);
    }
    const int64_t size1 = 1LL << n1, size2 = 1LL << n2;
    vector<int> arr1(size1), arr2(size2);
    for (auto &v : arr1) v = (int)(splitmix64(seed) % (uint64_t)alphabet);
    for (auto &v : arr2) v = (int)(splitmix64(seed) % (uint64_t)alphabet);
    mkdir("./Data", 0755);
    while (n1 >= WF_GEN_MIN_EXP && n2 >= 0) {
        const int64_t s1 = 1LL << n1, s2 = 1LL << n2;
        const string path = "./Data/str1_2_" + to_string(n1) + "_str2_2_" + to_string(n2) + ".txt";
        ofstream f(path.c_str());
        f << s1 << " " << s2 << endl;
        for (int64_t i = 0; i < s1; i++) f << arr1[i] << " ";
        f << endl;
        for (int64_t i = 0; i < s2; i++) f << arr2[i] << " ";
        f << endl;
        n1--;
        n2--;
    }
    return 0;
}
