/* SwiftPM needs one source file in a C target; the code lives in libstrouhal_hip.so. */
