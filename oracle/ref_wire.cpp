// Golden-vector generator for the wire formats -- TEST INFRASTRUCTURE ONLY.
// Compiled by oracle/Makefile (target ref) against the reference's own
// headers (src/Networking/onlineSpikesPayload.h, sorterParameters.h) and its
// vendored cereal (External/cereal); writes the cereal binary serialisation
// of fixed payloads to the directory given as argv[1].  This build is LP64:
// the structs' `long` fields are 8 bytes here.
#include <cstdint>
#include <cstring>   // FragmentManager.h uses memcpy without including it
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "onlineSpikesPayload.h"
#include "sorterParameters.h"

template <class T>
static void dump(const T& v, const std::string& path) {
    std::stringstream ss;
    {
        cereal::BinaryOutputArchive ar(ss);
        ar(const_cast<T&>(v));
    }
    const std::string s = ss.str();
    std::ofstream(path, std::ios::binary).write(s.data(), (std::streamsize)s.size());
    std::printf("%s %zu bytes\n", path.c_str(), s.size());
}

int main(int argc, char** argv) {
    const std::string out = argc > 1 ? argv[1] : ".";
    OnlineSpikesPayload p{};
    p.recordingOffset = 123456789;
    p.streamSampleCt = 9876543210L;
    p.Times = {1000, 1003, 2047, 3000000001L};
    p.Templates = {5, 17, 5, 1023};
    p.Amplitudes = {12.5f, -3.25f, 1e-3f, 65504.0f};
    p.VRMS = 7.125;
    p.P2P = 812.0f;
    p.processTime = 431;
    p.eventStreamSampleCt = -7;
    p.predictLabel = 3;
    p.label = -2;
    p.nTrials = 40;
    p.nCorrect = 31;
    p.confidence = 0.8125;
    dump(p, out + "/payload_lp64.bin");
    OnlineSpikesPayload e{};
    dump(e, out + "/payload_empty_lp64.bin");
    SorterParameters s{};
    s.m_lT = 1024;
    s.m_lC = 384;
    s.m_lM = 61;
    s.m_lN = 3122;
    s.m_fSampRate = 30000.0f;
    s.m_dChanpos = {0.0, 20.0, 43.5, -16.25};
    s.m_vNeuronIndices = {0, 1, 2, 7, 1023};
    dump(s, out + "/sorter_params_lp64.bin");
    return 0;
}
