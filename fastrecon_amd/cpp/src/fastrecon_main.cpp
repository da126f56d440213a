// The recon program (reference Program/Main.cpp:98-197 and
// Program/ProgramOptions.cpp:9-134): RECON.ini + options -> gridding ->
// FFT -> fftshift -> SOS (or PILS) -> |image| as float32 to out_file.
// Qt is replaced by argp, a small INI reader and a name-sorted glob; the
// --show window is not built (it needs a display).
#include <argp.h>
#include <dirent.h>
#include <fnmatch.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <string>
#include <vector>

#include "GridLut.h"
#include "ImageFilter.h"
#include "ReconData.h"

namespace {

struct ReconParameters {          // ProgramOptions.h:6-23
    std::string path;
    std::string trajFiles;
    std::string dcfFile;
    std::string dataFiles;
    std::string outFile;
    int samples = 0;
    int projections = 0;
    float overgridding_factor = 0;
    float kernel_width = 4;
    bool pils = false;
    unsigned rcxres = 0, rcyres = 0, rczres = 0;
};

struct Options {
    ReconParameters params;
    std::string iniFileName;
    int arg_count = 1;
    bool display = false;
    bool gpu = false;
};

// ProgramOptions.cpp:82-119
error_t parse_opt(int key, char *arg, struct argp_state *state)
{
    auto *o = static_cast<Options *>(state->input);
    switch (key) {
    case 666: o->gpu = true; break;
    case 777: o->display = true; break;
    case 888: o->params.pils = true; break;
    case 'g': o->params.overgridding_factor = atof(arg); break;
    case 'w': o->params.kernel_width = atof(arg); break;
    case 'h': argp_usage(state); break;
    case ARGP_KEY_ARG:
        o->arg_count--;
        if (o->arg_count >= 0)
            o->iniFileName = arg;
        break;
    case ARGP_KEY_END:
        if (o->arg_count >= 1)
            argp_failure(state, 1, 0, "Missing RECON.ini arguments");
        else if (o->arg_count < 0)
            argp_failure(state, 1, 0, "too many arguments");
        break;
    }
    return 0;
}

std::string trim(const std::string &s)
{
    auto b = s.find_first_not_of(" \t\r\n");
    if (b == std::string::npos) return "";
    auto e = s.find_last_not_of(" \t\r\n");
    std::string t = s.substr(b, e - b + 1);
    if (t.size() >= 2 && t.front() == '"' && t.back() == '"') t = t.substr(1, t.size() - 2);
    return t;
}

// QSettings(IniFormat) keys without a group (ProgramOptions.cpp:15-46).
std::map<std::string, std::string> readIni(const std::string &name)
{
    std::map<std::string, std::string> kv;
    std::ifstream f(name);
    std::string line, group;
    while (std::getline(f, line)) {
        std::string t = trim(line);
        if (t.empty() || t[0] == ';' || t[0] == '#') continue;
        if (t[0] == '[') {
            group = t.substr(1, t.find(']') - 1);
            continue;
        }
        auto eq = t.find('=');
        if (eq == std::string::npos) continue;
        if (!group.empty() && group != "General") continue;
        kv[trim(t.substr(0, eq))] = trim(t.substr(eq + 1));
    }
    return kv;
}

int toInt(const std::map<std::string, std::string> &kv, const char *k)
{
    auto it = kv.find(k);
    return it == kv.end() ? 0 : std::atoi(it->second.c_str());
}

void loadIniOptions(Options &o)
{
    std::ifstream test(o.iniFileName);
    if (!test.good()) {
        std::cerr << "File '" << o.iniFileName << "' does not exist!" << std::endl;   // ProgramOptions.cpp:17-22
        exit(1);
    }
    auto slash = o.iniFileName.find_last_of('/');
    o.params.path = slash == std::string::npos ? "" : o.iniFileName.substr(0, slash + 1);
    auto kv = readIni(o.iniFileName);
    o.params.trajFiles = kv["traj_file"];
    o.params.dcfFile = kv["weight_file"];
    o.params.dataFiles = kv["data_file"];
    o.params.outFile = kv["out_file"];
    o.params.samples = toInt(kv, "xres");
    o.params.projections = toInt(kv, "projections");
    o.params.rcxres = toInt(kv, "rcxres");
    o.params.rcyres = toInt(kv, "rcyres");
    o.params.rczres = toInt(kv, "rczres");
    if (o.params.overgridding_factor == 0)
        o.params.overgridding_factor = std::atof(kv["overgridding_factor"].c_str());
}

void showParameters(const ReconParameters &p)      // ProgramOptions.cpp:121-129
{
    std::cout << std::left << std::setw(22) << "Samples:" << p.samples << std::endl
              << std::setw(22) << "Projections:" << p.projections << std::endl
              << std::setw(22) << "Overgridding factor:" << p.overgridding_factor << std::endl
              << std::setw(22) << "Kernel width:" << p.kernel_width << std::endl
              << std::setw(22) << "rcxres:" << p.rcxres << std::endl
              << std::setw(22) << "rcyres:" << p.rcyres << std::endl
              << std::setw(22) << "rczres:" << p.rczres << std::endl;
}

// QDir(path, pattern, QDir::Name).entryList(), prefixed with path.
std::vector<std::string> globSorted(const std::string &path, const std::string &pattern)
{
    std::vector<std::string> names;
    if (pattern.empty()) return names;
    DIR *d = opendir(path.empty() ? "." : path.c_str());
    if (!d) return names;
    while (dirent *e = readdir(d)) {
        std::string n = e->d_name;
        if (n == "." || n == "..") continue;
        if (fnmatch(pattern.c_str(), n.c_str(), 0) == 0) names.push_back(n);
    }
    closedir(d);
    std::sort(names.begin(), names.end());
    for (auto &n : names) n = path + n;
    return names;
}

class Timer {
public:
    Timer() { restart(); }
    long long elapsed() const {
        return std::chrono::duration_cast<std::chrono::milliseconds>(Clock::now() - t0).count();
    }
    long long restart() { auto e = elapsed(); t0 = Clock::now(); return e; }
private:
    using Clock = std::chrono::steady_clock;
    Clock::time_point t0;
};

}  // namespace

int main(int argc, char *argv[])
{
    Options options;
    struct argp_option opts[] = {
        {0, 0, 0, 0, "Recon parameters:", 1},
        {0, 'g', "FACTOR", 0, "Over-gridding factor", 0},
        {0, 'w', "WIDTH", 0, "Kernel full width", 0},
        {0, 0, 0, 0, "Recon options:", 2},
        {"pils", 888, 0, 0, "PILS method", 0},
        {0, 0, 0, 0, "Miscellaneous:", 3},
        {"show", 777, 0, 0, "Display result in a window (not available in this build)", 0},
        {"gpu", 666, 0, 0, "Use GPU computation (always on in this build)", 0},
        {0, 0, 0, 0, "Help options:", -1},
        {0, 'h', 0, OPTION_HIDDEN, 0, 0},
        {0, 0, 0, 0, 0, 0}};
    struct argp argp = {opts, parse_opt, "RECON.ini",
                        "MI355X MRI recon program.\v"
                        "Command line options have precedence over RECON.ini.",
                        0, 0, 0};
    argp_parse(&argp, argc, argv, 0, 0, &options);
    loadIniOptions(options);
    showParameters(options.params);
    const ReconParameters &params = options.params;

    // -------------- Load multi-channel data (Main.cpp:79-96) --------
    auto reconData = ReconData<float>::Create(params.samples, params.projections, options.gpu);
    reconData->loadFromFiles(globSorted(params.path, params.dataFiles), globSorted(params.path, params.trajFiles),
                             params.path + params.dcfFile);

    Timer timer0, timer;

    // -------------- Gridding ---------------------------------------
    auto grid = GridLut<float>::Create(*reconData);
    grid->plan(params.rcxres, params.overgridding_factor, params.kernel_width, 512);
    auto imgData = grid->execute();
    std::cout << "Gridding total time " << timer.elapsed() << " ms" << std::endl;

    ImageData<float> imgMap;
    if (params.pils)
        imgMap = *imgData;

    auto filter = ImageFilter<float>::Create(*imgData);

    // --------------- FFT -------------------------------------------
    filter->fftPlan();
    timer.restart();
    filter->fftExecute();
    filter->fftShift();
    std::cout << "FFT total time " << timer.restart() << " ms" << std::endl;

    // -------------- Recon methods ----------------------------------
    timer.restart();
    if (params.pils) {
        auto filterMap = ImageFilter<float>::Create(imgMap);
        std::cout << "\nRecon PILS... " << std::endl;
        filterMap->lowFilter(22);
        std::cout << "\nLow pass filtering | " << timer.restart() << " ms" << std::endl;
        std::cout << "\nFFT low res image... " << std::endl;
        filterMap->fftExecute();
        filterMap->fftShift();
        std::cout << "FFT total time " << timer.restart() << " ms" << std::endl;
        filterMap->normalize();
        std::cout << "\nSum of Square Field Map..." << std::flush;
        filter->SOS(imgMap, {params.rcxres, params.rcyres, params.rczres});
        std::cout << " | " << timer.elapsed() << " ms" << std::endl;
    } else {
        std::cout << "\nRecon SOS... " << std::flush;
        filter->SOS({params.rcxres, params.rcyres, params.rczres});
        std::cout << " | " << timer.elapsed() << " ms" << std::endl;
    }
    std::cout << "\nProgram total time excluding I/O: " << timer0.elapsed() / 1000.0 << " s" << std::endl;

    // -------------- Save (Main.cpp:179-186) ---------------------------
    std::ofstream file(params.path + params.outFile, std::ios::binary);
    for (const auto &data : *imgData->getChannelImage()) {
        float value = std::abs(data);
        file.write(reinterpret_cast<const char *>(&value), sizeof(value));
    }
    file.close();

    if (options.display)
        std::cerr << "--show: no display support in this build; the image was written to "
                  << params.path + params.outFile << std::endl;
    return 0;
}
