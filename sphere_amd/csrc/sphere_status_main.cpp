// sphere_status_main.cpp -- the `sphere_status` tool (src/sphere_status.cpp:
// 4-47): reads output/<sid>.status.dat, which startTime writes as
// "%2.4e %2.4e %d" (time, percent complete, last output file number), and
// prints it.  The latest output file is named as startTime names it
// (<sid>.outputNNNNN.bin, five digits); the reference prints the number
// without the padding.
#include <unistd.h>

#include <cstdio>
#include <iostream>
#include <string>

int main(int argc, char* argv[]) {
    if (argc != 2) {
        std::cout << "You need to specify the simulation ID as an input parameter,\n"
                  << "e.g. " << argv[0] << " 3dtest\n";
        return 1;
    }
    char* cwd = getcwd(nullptr, 0);
    if (!cwd) {
        std::cout << "getcwd failed\n";
        return 1;
    }
    const std::string file = std::string(cwd) + "/output/" + argv[1] + ".status.dat";
    free(cwd);
    FILE* fp = std::fopen(file.c_str(), "r");
    if (!fp) {
        std::cerr << argv[0] << ": cannot open " << file << "\n";
        return 1;
    }
    float time_current = 0.0f, percent = 0.0f;
    int file_nr = 0;
    const int got = std::fscanf(fp, "%f%f%d", &time_current, &percent, &file_nr);
    std::fclose(fp);
    if (got != 3) {
        std::cerr << argv[0] << ": " << file << " is not a status file\n";
        return 1;
    }
    char name[32];
    std::snprintf(name, sizeof name, ".output%05d.bin", file_nr);
    std::cout << "Reading " << file << ":\n"
              << " - Current simulation time:  " << time_current << " s\n"
              << " - Percentage completed:     " << percent << "%\n"
              << " - Latest output file:       " << argv[1] << name << "\n";
    return 0;
}
