// xsec_main.cpp -- gexs_xsec: keyword input on stdin -> cross-sections on stdout.
//
// The flow of the reference's main.cpp:17-43 (read input, build the profile
// driver, Initialize, ExecuteCrossSection, OutputProfile), with HipManager as
// the backend.  Exit status 1 on any error (message on stderr).
#include <cstdio>
#include <iostream>

#include "gexs/CrossSection.h"

int main() {
    try {
        gexs::Input in = gexs::ParseInput(std::cin);
        gexs::CrossSection xs(in);
        xs.Initialize();
        xs.ExecuteCrossSection();
        xs.OutputProfile(stdout);
        std::fprintf(stderr, "lines read %lld, selected %lld, chunks %d\n", (long long)xs.LinesRead(),
                     (long long)xs.LinesSelected(), xs.Chunks());
    } catch (const std::exception& e) {
        std::fprintf(stderr, "gexs_xsec: %s\n", e.what());
        return 1;
    }
    return 0;
}
