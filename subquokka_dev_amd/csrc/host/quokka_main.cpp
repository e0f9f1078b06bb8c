// quokka -- command line with the reference's interface: quokka -i <ini> -c <circuit>
// (main.cpp:13-20, readargs common.cpp:5-32).  Errors print the reference's
// message and exit(1).
#include <cstdlib>
#include <iostream>
#include <string>

#include "simulator.hpp"

namespace {

void readargs(int argc, char *argv[], std::string &ini, std::string &circuit) {
    for (int i = 1; i < argc; ++i) {
        std::string arg(argv[i]);
        if (arg.length() > 1 && arg[0] == '-') {
            std::string option = arg.substr(1);
            if ((option == "i" || option == "c") && i + 1 >= argc) {
                std::cout << "[Missing Value]: -" << option << std::endl;
                std::exit(1);
            }
            if (option == "i") {
                ini = argv[++i];
                std::cout << "[Config File]: " << ini << std::endl;
            } else if (option == "c") {
                circuit = argv[++i];
                std::cout << "[Circuit File]: " << circuit << std::endl;
            } else {
                std::cout << "[Unknown Option]: " << option << std::endl;
                std::exit(1);
            }
        } else {
            std::cout << "[Non-option Argument]: " << arg << std::endl;
            std::exit(1);
        }
    }
}

}  // namespace

int main(int argc, char *argv[]) {
    std::string ini, cir;
    readargs(argc, argv, ini, cir);
    try {
        sq::Simulator quokka(ini, cir);
        quokka.run();
    } catch (const sq::SimError &e) {
        std::cerr << e.what() << std::endl;
        return 1;
    }
    return 0;
}
