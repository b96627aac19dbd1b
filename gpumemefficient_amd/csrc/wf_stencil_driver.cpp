// wf_stencil_driver.cpp -- drop-in drivers of the reference's SAT and 2D-SOR
// hyperlane programs, with their argv and stdout.
//
// One source, compiled per binary (gpumemefficient_amd/build.py), linked with
// libwfdp_compat_sat.so or libwfdp_compat_sor.so (both export SOR()):
//   -DWF_SDRIVER_SHARE5  SAT/GPU_Hyperlane_Share.cpp:44-100,
//                        2D-SOR/nested_loop/GPU_Hyperlane_Share.cpp:44-100:
//                        argv "e1 e2 tX tY itr"
//   (default)            the Cache drivers (SAT/GPU_Hyperlane_Cache.cpp:44-95):
//                        argv "e1 e2", itr = 100
// The input is ./Data/x_2_<e1>_y_2_<e2>.txt: "n1 n2", then n2 + 3 rows of
// n1 + 3 values (SAT/GPU_Hyperlane_Share.cpp:12-30).  The timed region is the
// loop of SOR() calls and the only output line is the mean time per call, as
// in the reference (the swept table is discarded there too).  With
// WFDP_DEBUG_TABLE=1 one more, untimed call writes the swept interior to
// ./Output/output_<Program>_<e1>.txt: first line n1, then n2 lines of n1
// space-terminated values (the reference's DEBUG block cannot compile: it
// prints an undeclared `table`, SAT/GPU_Hyperlane_Share.cpp:103-121).
#include <sys/time.h>

#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "wfdp_compat.h"

#ifndef WF_SDRIVER_NAME
#define WF_SDRIVER_NAME "SOR"
#endif

using namespace std;

static void readInputData(const string &path, int &n1, int &n2, vector<int> &arr)
{
    ifstream in(path.c_str());
    if (!in) {
        cout << "ERROR: Input file cannot be opened!" << endl;
        exit(EXIT_FAILURE);
    }
    in >> n1 >> n2;
    if (n1 < 0 || n2 < 0) {
        cout << "ERROR: Input file cannot be opened!" << endl;
        exit(EXIT_FAILURE);
    }
    arr.assign((size_t)(n1 + 3) * (size_t)(n2 + 3), 0);
    for (size_t i = 0; i < arr.size(); i++) in >> arr[i];
}

int main(int argc, char **argv)
{
    int nn1 = 0, nn2 = 0, tX = 0, tY = 0, itr = 100;
#if defined(WF_SDRIVER_SHARE5)
    if (argc != 6) {
        // SAT/GPU_Hyperlane_Share.cpp:46 and
        // 2D-SOR/nested_loop/GPU_Hyperlane_Share.cpp:46 word this line
        // differently.
        if (std::string(WF_SDRIVER_NAME) == "SAT")
            cout << "Incorrect Input Parameters. Must be two string sizes, two tile sizes, and one iteration step."
                 << endl;
        else
            cout << "Incorrect Input Parameters. Must be two string sizes, two tile sizes, and one total iteration "
                    "step."
                 << endl;
        exit(EXIT_FAILURE);
    }
    nn1 = atoi(argv[1]);
    nn2 = atoi(argv[2]);
    tX = atoi(argv[3]);
    tY = atoi(argv[4]);
    itr = atoi(argv[5]);
#else
    if (argc != 3) {
        cout << "Incorrect Input Parameters. Must be two string sizes." << endl;
        exit(EXIT_FAILURE);
    }
    nn1 = atoi(argv[1]);
    nn2 = atoi(argv[2]);
#endif
    if (itr < 1) itr = 1;
    ostringstream path;
    path << "./Data/x_2_" << nn1 << "_y_2_" << nn2 << ".txt";
    int n1 = 0, n2 = 0;
    vector<int> arr;
    readInputData(path.str(), n1, n2, arr);

    const char *dbg = getenv("WFDP_DEBUG_TABLE");
    const bool dump = dbg && *dbg && *dbg != '0';
    vector<int> orig;
    if (dump) orig = arr;  // the timed calls write back too under WFDP_DEBUG_TABLE

    timeval tbegin, tend;
    gettimeofday(&tbegin, NULL);
    for (int i = 0; i < itr; i++) {
#if defined(WF_SDRIVER_SHARE5)
        SOR(n1, n2, arr.data(), tX, tY);
#else
        SOR(n1, n2, arr.data());
#endif
    }
    gettimeofday(&tend, NULL);
    const double ms = (double)(tend.tv_sec - tbegin.tv_sec) * 1000.0 + (double)(tend.tv_usec - tbegin.tv_usec) / 1000.0;
    cout << "execution time: " << ms / itr << " ms." << endl;

    if (dump) {
        // One sweep of the input; WFDP_DEBUG_TABLE makes it write the result back.
        arr = orig;
        SOR(n1, n2, arr.data(), tX, tY);
        ostringstream outfile;
        outfile << "./Output/output_" << WF_SDRIVER_NAME << "_" << nn1 << ".txt";
        ofstream output(outfile.str().c_str());
        if (!output) {
            cout << "ERROR: output file cannot be opened!" << endl;
            exit(EXIT_FAILURE);
        }
        output << n1 << '\n';
        for (int y = 1; y <= n2; y++) {
            for (int x = 1; x <= n1; x++) output << arr[(size_t)y * (n1 + 3) + x] << ' ';
            output << '\n';
        }
    }
    return 0;
}
