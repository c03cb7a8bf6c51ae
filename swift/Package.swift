// swift-tools-version:5.9
// SwiftPM binding of libstrouhal_hip.so (the C ABI in ../include/strouhal_hip.h).
// Build the library first (python -c 'import __graft_entry__ as g; g.build()'),
// then point STROUHAL_HIP_LIB at the directory that holds libstrouhal_hip.so.
import PackageDescription
import Foundation

let libDir = ProcessInfo.processInfo.environment["STROUHAL_HIP_LIB"] ?? "../strouhal_amd"

let package = Package(
    name: "StrouhalHIP",
    products: [.library(name: "StrouhalCore", targets: ["StrouhalCore"])],
    targets: [
        .target(name: "CStrouhalHIP",
                linkerSettings: [.unsafeFlags(["-L\(libDir)", "-Xlinker", "-rpath", "-Xlinker", libDir]),
                                 .linkedLibrary("strouhal_hip")]),
        .target(name: "StrouhalCore", dependencies: ["CStrouhalHIP"]),
    ]
)
