// xsec_main.cpp -- gexs_xsec: keyword input on stdin -> cross-sections on stdout.
//
// The flow of the reference's main.cpp:17-43 (read input, build the profile
// driver, Initialize, ExecuteCrossSection, OutputProfile), with HipManager as
// the backend.  Exit status 1 on any error (message on stderr).
//
// gexs_xsec --lines: read the input's line files through the same reader and
// print every line as "nu g' E'' A gamma n" (%.17g, round-trip exact), in
// file order; no device is used.
#include <cstdio>
#include <cstring>
#include <iostream>

#include "gexs/CrossSection.h"

static int dump_lines(const gexs::Input& in) {
    std::unique_ptr<gexs::StateReader> r = gexs::MakeReader(in);
    double nu, gns, e_i, aif, gam, n;
    for (const std::string& f : in.trans_files) {
        r->OpenFile(f);
        while (r->ReadNextState(nu, gns, e_i, aif, gam, n))
            std::printf("%.17g %.17g %.17g %.17g %.17g %.17g\n", nu, gns, e_i, aif, gam, n);
        r->CloseFile();
    }
    return 0;
}

int main(int argc, char** argv) {
    try {
        gexs::Input in = gexs::ParseInput(std::cin);
        if (argc > 1 && std::strcmp(argv[1], "--lines") == 0) return dump_lines(in);
        gexs::CrossSection xs(in);
        xs.Initialize();
        xs.ExecuteCrossSection();
        xs.OutputProfile(stdout);
        std::fprintf(stderr, "lines read %lld, selected %lld, chunks %d\n", (long long)xs.LinesRead(),
                     (long long)xs.LinesSelected(), xs.Chunks());
        // the two band-strength sums of the report line at full precision
        std::fprintf(stderr, "band strength: summed %.17g integrated %.17g\n", xs.TotalIntensity(),
                     xs.IntegratedIntensity());
    } catch (const std::exception& e) {
        std::fprintf(stderr, "gexs_xsec: %s\n", e.what());
        return 1;
    }
    return 0;
}
