// Compile-and-link check of the reference-side adapter include/atjup_hip_step.hpp against the
// reference's own headers and objects (oracle/Makefile, target `adapter`). The model header it sees
// is the reference's planet/cJupiterModel.h with the one friend line the adapter asks a maintainer to
// add; every member function of HipStep<cJupiterModel> is instantiated, so each model member and
// library call it names is checked, and the binary is linked against libatjup_hip.so. It is not run:
// the model's start-up (Run(), planet/cJupiterModel.cpp:316-745) is not separable from its loop.
#include "cJupiterModel.h"
#include "atjup_hip_step.hpp"

#include <cstdio>

template class HipStep<cJupiterModel>;

int main(){
    std::printf("HipStep<cJupiterModel>: compiled against the reference headers and linked with libatjup_hip.so\n");
    return 0;
}
