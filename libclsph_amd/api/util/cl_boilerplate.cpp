// readKernelFile: util/cl_boilerplate.cpp:90-112.
#include "cl_boilerplate.h"

#include <cstdlib>
#include <fstream>
#include <iostream>
#include <iterator>

std::string readKernelFile(std::string kernelFileName) {
  std::ifstream in(kernelFileName, std::ios::in | std::ios::binary);
  if (!in) {
    std::cout << kernelFileName << " file not found." << std::endl;
    std::exit(1);
  }
  std::string text((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
  if (text.empty()) {
    std::cout << kernelFileName << " is empty." << std::endl;
    std::exit(1);
  }
  return text;
}
