// main_axion_strings.cpp -- command-line driver (the reference's
// Main_AxionStrings.cpp:10-80 run loop, single level).
//
//   axion_strings params.txt [key=value ...] [--check] [--out DIR]
//
// --check only parses and validates the parameters (the reference's
// just_check_params), without touching a GPU, and lists the keys the box
// plan derived ("derived key = values").
#include "AxionStringsLevel.hpp"

#include <chrono>
#include <cstring>
#include <iostream>

int main(int argc, char **argv)
{
    using namespace axs_host;
    if (argc < 2)
    {
        std::cerr << "usage: " << argv[0]
                  << " params.txt [key=value ...] [--check] [--out DIR]\n";
        return 2;
    }
    ParamTable table;
    bool check_only = false;
    std::string out_dir = ".";
    try
    {
        table.read_file(argv[1]);
        // a relative tuning table named in the file is relative to the file
        // (one given on the command line, to the working directory)
        const std::string tkey = "axion_strings.tuning_table";
        if (table.contains(tkey))
        {
            const std::string tp = table.values(tkey).front();
            const std::string pf = argv[1];
            const size_t slash = pf.find_last_of('/');
            if (!tp.empty() && tp[0] != '/' && slash != std::string::npos)
                table.set(tkey, pf.substr(0, slash + 1) + tp);
        }
        for (int i = 2; i < argc; ++i)
        {
            if (!std::strcmp(argv[i], "--check"))
                check_only = true;
            else if (!std::strcmp(argv[i], "--out") && i + 1 < argc)
                out_dir = argv[++i];
            else
                table.add_line(argv[i]);
        }
        const ParamTable given = table;
        Params::check_params(table, &std::cout);
        if (check_only)
        {
            // the keys the box plan added, as the run would read them
            for (const auto &kv : table.all())
                if (!given.contains(kv.first))
                {
                    std::cout << "derived " << kv.first << " =";
                    for (const auto &v : kv.second)
                        std::cout << " " << v;
                    std::cout << "\n";
                }
            std::cout << "parameters ok\n";
            return 0;
        }
        using clock = std::chrono::steady_clock;
        const auto t0 = clock::now();
        AxionStringsLevel level(table, out_dir);
        const auto t1 = clock::now();
        level.initData();
        const auto t2 = clock::now();
        level.evolve();
        const auto t3 = clock::now();
        auto sec = [](clock::duration d) {
            return std::chrono::duration<double>(d).count();
        };
        std::cout << "done: tau = " << level.tau() << "  steps = "
                  << level.step() << "\n";
        std::cout << "wall (s): set-up " << sec(t1 - t0) << "  initial data "
                  << sec(t2 - t1) << "  evolution " << sec(t3 - t2)
                  << " (of which diagnostics " << level.diagnostics_seconds()
                  << ")\n";
    }
    catch (const ParamError &e)
    {
        std::cerr << "parameter error: " << e.what() << "\n";
        return 3;
    }
    catch (const std::exception &e)
    {
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
