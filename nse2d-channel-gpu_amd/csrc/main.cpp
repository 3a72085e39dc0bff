// nse2d / nse2d_freeslip executables: the reference's command line (init.cu:7-96). Invoked under
// the name nse2d_freeslip the wall closure is free-slip, like the reference's WALLS=freeslip build.
#include "../../include/nse2d_hip.h"

#include <cstring>
#include <string>
#include <vector>

int main(int argc, char** argv) {
    std::vector<char*> args(argv + 1, argv + argc);
    static char walls[] = "--walls", freeslip[] = "freeslip";
    const std::string prog(argv[0]);
    const std::string tag = "nse2d_freeslip";
    if (prog.size() >= tag.size() && prog.compare(prog.size() - tag.size(), tag.size(), tag) == 0) {
        args.push_back(walls);
        args.push_back(freeslip);
    }
    return nse2d_main((int)args.size(), args.data());
}
