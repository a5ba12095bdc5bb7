// lmultislice -- the LMultiSlice command-line program (LMultiSlice.cpp:6-21)
// over the msx engine: runs the XML command of Dispatcher::Run
// (Dispatcher.cu:161-242) through the C++ mirror in msx/lmultislice.hpp.
//
//   lmultislice CONFIG.xml          run: <output>_pt.mrc (potential) and
//                                   <output>.mrc (|psi| of the image wave)
//   lmultislice                     asks for the XML file name on stdin
//   lmultislice --parse CONFIG.xml  print the parsed command (no GPU work)
//   lmultislice --read-coo FILE     print the model read from a .coo file
//                                   (no GPU work)
//
// The exit status is the run's return value, as the reference's main
// (0, or -1 -> 255).
#include <cstdio>
#include <cstring>
#include <iostream>
#include <string>

#include "msx/lmultislice.hpp"

using namespace msx::lms;

int main(int argc, char* argv[]) {
    if (argc >= 3 && std::strcmp(argv[1], "--parse") == 0) {
        Command c;
        if (parse_command(argv[2], c) == -1) return -1;
        std::printf("fileNameInput=%s\nfileNameOutput=%s\nnx=%d\nny=%d\ndpa=%.9g\nradiuc=%.9g\nnumberslices=%d\n"
                    "aperture=%.9g\ncs=%.9g\ndefocus=%.9g\nkeV=%.9g\n",
                    c.fileNameInput.c_str(), c.fileNameOutput.c_str(), c.nx, c.ny, c.dpa, c.radiuc, c.numberSlices,
                    c.aperture, c.cs, c.defocus, c.keV);
        return 0;
    }
    if (argc >= 3 && std::strcmp(argv[1], "--read-coo") == 0) {
        amodel::Model m;
        if (m.read(argv[2]) == -1) return -1;
        std::printf("cell=%.9g %.9g %.9g %.9g %.9g %.9g\natoms=%zu\n", m.getA(), m.getB(), m.getC(), m.getAlpha(),
                    m.getBeta(), m.getGamma(), m.getNumberAtoms());
        const amodel::Cortege* t = m.getTableCell();
        for (size_t i = 0; i < m.getNumberAtoms(); i++) {
            const auto& e = t[i].element;
            std::printf("%s %.9g %.9g %.9g %.9g %.9g %.9g\n", e.Atom, e.xsCoordinate.x, e.xsCoordinate.y,
                        e.xsCoordinate.z, e.absorbtion, e.oxygenation, e.occupation);
        }
        return 0;
    }
    std::string xml;
    if (argc == 1) {
        std::cout << "Input name XML file: " << std::endl;
        std::cin >> xml;
    } else {
        std::cout << "Name XML file: " << argv[1] << std::endl;
        xml = argv[1];
    }
    return dispatcher_run(xml.c_str());
}
