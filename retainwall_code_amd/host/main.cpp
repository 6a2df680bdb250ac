// main.cpp -- `loquat <project> [gpu_id]` (main.cpp:23-29 of the reference).
#include <iostream>

#include "loquat.hpp"

int main(int argc, char* argv[]) {
  if (argc < 2) {
    std::cout << "Error! Please give the project name!" << std::endl;
    return 1;
  }
  LoquatSph sph(argv[1]);
  sph.RunSimulation(argc, argv);
  return 0;
}
