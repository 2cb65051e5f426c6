// racer: the command-line driver of main.mojo:1-149 over the C ABI.
//
//   racer train     TRACK [auto|gpu|cpu] [cars] [iterations] [checkpoint.npy] [resume.npy]
//   racer eval      TRACK [auto|gpu|cpu] [cars] [steps] [checkpoint.npy]
//   racer benchmark TRACK [auto|gpu|cpu] [cars] [steps]
//   racer serve     TRACK [auto|gpu|cpu] [cars] [port] [checkpoint.npy|-]
//
// TRACK is a built-in name (flat, ramp, bank, bridge, ring10, ramp10, bank10,
// bridge8, random<k>), a map YAML, or a float32 .npy track array. As the
// reference does (main.mojo:133, through Python), names and YAML maps are
// compiled by the Python track compiler, here in a child process
// (python3 -m warporacer3d_amd.track TRACK FILE.npy) run before the GPU is
// touched; `serve` is the Python server (python3 -m warporacer3d_amd.serve),
// also a child. Shape parameters come from the environment: RACER_BEAMS (108),
// RACER_HIDDEN (64), RACER_ROLLOUT (32), RACER_EPOCHS (4), RACER_MINIBATCHES
// (4), RACER_PRECISION (fp32|bf16); RACER_PYTHON names the interpreter.
// `cpu` fails: the library is the GPU path only.
#include <sys/wait.h>
#include <unistd.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "racer.h"

static const char* USAGE =
    "racer train     TRACK [auto|gpu|cpu] [cars] [iterations] [checkpoint.npy] [resume.npy]\n"
    "racer eval      TRACK [auto|gpu|cpu] [cars] [steps] [checkpoint.npy]\n"
    "racer benchmark TRACK [auto|gpu|cpu] [cars] [steps]\n"
    "racer serve     TRACK [auto|gpu|cpu] [cars] [port] [checkpoint.npy|-]\n"
    "TRACK is flat, ramp, bank, bridge, ring10, ramp10, bank10, bridge8, random<k>, a map YAML, or a .npy track "
    "array.";

// The directory holding the package (this binary sits in it), for PYTHONPATH.
static std::string package_parent() {
    char buf[4096];
    ssize_t len = readlink("/proc/self/exe", buf, sizeof(buf) - 1);
    if (len <= 0) return ".";
    std::string exe(buf, (size_t)len);
    std::string dir = exe.substr(0, exe.rfind('/'));  // .../warporacer3d_amd
    return dir.substr(0, dir.rfind('/'));
}

// Runs python3 -m <module> <args...> as a child (fork + exec in the child,
// before this process has touched the GPU); returns its exit status.
static int run_python(const std::string& module, const std::vector<std::string>& args) {
    const char* py = std::getenv("RACER_PYTHON");
    std::string python = py && *py ? py : "python3";
    std::string path = package_parent();
    const char* old = std::getenv("PYTHONPATH");
    std::string pp = old && *old ? path + ":" + old : path;
    pid_t pid = fork();
    if (pid < 0) return 127;
    if (pid == 0) {
        setenv("PYTHONPATH", pp.c_str(), 1);
        std::vector<char*> av;
        av.push_back(const_cast<char*>(python.c_str()));
        av.push_back(const_cast<char*>("-m"));
        av.push_back(const_cast<char*>(module.c_str()));
        for (const auto& a : args) av.push_back(const_cast<char*>(a.c_str()));
        av.push_back(nullptr);
        execvp(av[0], av.data());
        _exit(127);
    }
    int status = 0;
    if (waitpid(pid, &status, 0) < 0) return 127;
    return WIFEXITED(status) ? WEXITSTATUS(status) : 128 + (WIFSIGNALED(status) ? WTERMSIG(status) : 0);
}

static int fail(const char* what) {
    std::fprintf(stderr, "error: %s: %s\n", what, racer_last_error());
    return 1;
}

static bool read_npy(const std::string& path, std::vector<float>& out, std::string& err) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return err = "cannot read " + path, false;
    char magic[8];
    f.read(magic, 8);
    if (!f || std::memcmp(magic, "\x93NUMPY", 6) != 0) return err = path + " is not a .npy file", false;
    uint32_t hl = 0;
    if (magic[6] == 1) {
        uint16_t h = 0;
        f.read(reinterpret_cast<char*>(&h), 2);
        hl = h;
    } else {
        f.read(reinterpret_cast<char*>(&hl), 4);
    }
    std::string header(hl, ' ');
    f.read(&header[0], hl);
    bool f4 = header.find("'<f4'") != std::string::npos;
    bool f8 = header.find("'<f8'") != std::string::npos;
    if (!f4 && !f8) return err = path + " must hold float32 or float64", false;
    size_t lp = header.find("'shape': ("), rp = header.find(')', lp);
    if (lp == std::string::npos || rp == std::string::npos) return err = path + " has no shape", false;
    size_t count = 1;
    std::stringstream ss(header.substr(lp + 10, rp - lp - 10));
    std::string item;
    while (std::getline(ss, item, ','))
        if (item.find_first_not_of(' ') != std::string::npos) count *= (size_t)std::stoull(item);
    out.resize(count);
    if (f4) {
        f.read(reinterpret_cast<char*>(out.data()), (std::streamsize)(count * 4));
    } else {
        std::vector<double> tmp(count);
        f.read(reinterpret_cast<char*>(tmp.data()), (std::streamsize)(count * 8));
        for (size_t i = 0; i < count; ++i) out[i] = (float)tmp[i];
    }
    if (!f) return err = path + " is truncated", false;
    return true;
}

static int env_int(const char* name, int def) {
    const char* v = std::getenv(name);
    return v && *v ? std::atoi(v) : def;
}

static std::string arg(int argc, char** argv, int i, const char* def) { return argc > i ? argv[i] : def; }

static bool ends_with(const std::string& s, const char* tail) {
    size_t n = std::strlen(tail);
    return s.size() >= n && s.compare(s.size() - n, n, tail) == 0;
}

// TRACK as a float32 array: a .npy file directly, anything else through the track compiler.
static bool load_track(const std::string& name, std::vector<float>& track, std::string& err) {
    if (ends_with(name, ".npy")) return read_npy(name, track, err);
    char tmpl[] = "/tmp/racer_track_XXXXXX.npy";
    int fd = mkstemps(tmpl, 4);
    if (fd < 0) return err = "cannot create a temporary file", false;
    close(fd);
    int rc = run_python("warporacer3d_amd.track", {name, tmpl});
    bool ok = rc == 0 && read_npy(tmpl, track, err);
    if (rc != 0) err = "the track compiler failed on " + name;
    unlink(tmpl);
    return ok;
}

int main(int argc, char** argv) {
    std::string mode = arg(argc, argv, 1, "");
    if (argc < 3 || (mode != "train" && mode != "eval" && mode != "benchmark" && mode != "serve")) {
        std::puts(USAGE);
        return argc < 3 ? 0 : 2;
    }
    if (mode == "serve") {  // main.mojo:94-124 and racer/serve.py, in Python over the same library
        std::vector<std::string> rest(argv + 2, argv + argc);
        return run_python("warporacer3d_amd.serve", rest);
    }
    std::string choice = arg(argc, argv, 3, "auto");
    if (choice == "cpu") {
        std::fprintf(stderr, "error: libracer_hip runs on the GPU only; the CPU device is not provided\n");
        return 1;
    }
    std::vector<float> track;
    std::string err;
    if (!load_track(argv[2], track, err)) {
        std::fprintf(stderr, "error: %s\n", err.c_str());
        return 1;
    }
    int cars = std::atoi(arg(argc, argv, 4, "1024").c_str());
    int count = std::atoi(arg(argc, argv, 5, "1000").c_str());
    std::string path = arg(argc, argv, 6, "agent.npy");
    racer_config cfg;
    racer_config_default(&cfg);
    cfg.beams = env_int("RACER_BEAMS", cfg.beams);
    cfg.hidden = env_int("RACER_HIDDEN", cfg.hidden);
    cfg.rollout = env_int("RACER_ROLLOUT", cfg.rollout);
    cfg.epochs = env_int("RACER_EPOCHS", cfg.epochs);
    cfg.minibatches = env_int("RACER_MINIBATCHES", cfg.minibatches);
    const char* prec = std::getenv("RACER_PRECISION");
    cfg.precision = prec && std::strcmp(prec, "bf16") == 0 ? 1 : 0;

    racer_device* dev = nullptr;
    if (racer_device_create(1, &dev)) return fail("device");
    std::printf("Device: %s\n", racer_device_name(dev));
    racer_sim* sim = nullptr;
    racer_policy* pol = nullptr;
    racer_acts* acts = nullptr;
    racer_buf* scratch = nullptr;
    if (racer_sim_create(dev, track.data(), track.size(), cars, 42, &cfg, &sim)) return fail("sim");
    if (racer_policy_create(dev, 42, &cfg, &pol)) return fail("policy");
    if (racer_acts_create(dev, cars, &cfg, &acts)) return fail("activations");
    if (racer_alloc(dev, 2 * (size_t)cars, 0.0f, &scratch)) return fail("alloc");
    if (racer_sim_reset(sim)) return fail("reset");

    if (mode == "train") {  // main.mojo:197-211
        racer_trainer* tr = nullptr;
        if (racer_trainer_create(dev, cars, &cfg, &tr)) return fail("trainer");
        std::string resume = arg(argc, argv, 7, "");
        int first = 0;
        if (!resume.empty() && racer_load(pol, resume.c_str(), &first)) return fail("load");
        auto start = std::chrono::steady_clock::now();
        for (int it = first; it < first + count; ++it) {
            float reward = 0.0f, kl = 0.0f;
            if (racer_trainer_rollout(tr, sim, pol, acts, it, &reward)) return fail("rollout");
            if (racer_trainer_update(tr, pol, &kl)) return fail("update");
            if (!std::isfinite(kl)) {
                std::fprintf(stderr, "error: training diverged (nonfinite KL)\n");
                return 1;
            }
            std::printf("iteration %d reward/step %g KL %g\n", it + 1, reward, kl);
            if ((it + 1) % 100 == 0 && racer_save(pol, path.c_str(), it + 1)) return fail("save");
        }
        if (racer_save(pol, path.c_str(), first + count)) return fail("save");
        double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
        std::printf("Saved %s ; %g training transitions/s\n", path.c_str(),
                    (double)cars * cfg.rollout * count / s);
        racer_trainer_destroy(tr);
    } else if (mode == "eval") {  // main.mojo:213-223
        if (racer_load(pol, path.c_str(), nullptr)) return fail("load");
        double reward = 0.0;
        long crashes = 0;
        std::vector<float> r(cars), d(cars);
        for (int t = 0; t < count; ++t) {
            if (racer_runner_step(sim, pol, acts, scratch, 1, t)) return fail("step");
            if (racer_read(racer_sim_reward(sim), r.data()) || racer_read(racer_sim_done(sim), d.data()))
                return fail("read");
            for (int i = 0; i < cars; ++i) {
                reward += r[i];
                crashes += d[i] == 1.0f;
            }
        }
        std::printf("reward/step %g crashes/car %g\n", reward / ((double)cars * count), (double)crashes / cars);
    } else {  // main.mojo:225-235
        double ms = 0.0;
        if (racer_benchmark(sim, pol, acts, scratch, count, &ms)) return fail("benchmark");
        std::printf("%d cars; %g ms/step; %g transitions/s\n", cars, ms, cars * 1000.0 / ms);
    }
    racer_free(scratch);
    racer_acts_destroy(acts);
    racer_policy_destroy(pol);
    racer_sim_destroy(sim);
    racer_device_destroy(dev);
    return 0;
}
