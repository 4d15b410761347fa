//! Compiles the engine's HIP sources (../../csrc/*.hip) for gfx950 with
//! hipcc, links them into OUT_DIR/libomrntt.so and tells cargo to link it.
//! The flags are those of omr-ntt_amd/build.py.  HIPCC overrides the
//! compiler path (default /opt/rocm/bin/hipcc); OMRNTT_JOBS the parallelism.
use std::path::{Path, PathBuf};
use std::process::Command;

fn main() {
    let manifest = PathBuf::from(std::env::var("CARGO_MANIFEST_DIR").unwrap());
    let csrc = manifest.join("../../csrc").canonicalize().expect("engine sources next to the crate");
    let out = PathBuf::from(std::env::var("OUT_DIR").unwrap());
    let hipcc = std::env::var("HIPCC").unwrap_or_else(|_| "/opt/rocm/bin/hipcc".into());
    let flags = [
        "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
        // packed FP32 instructions stay off
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    ];
    let mut srcs: Vec<PathBuf> = std::fs::read_dir(&csrc)
        .unwrap()
        .map(|e| e.unwrap().path())
        .filter(|p| p.extension().map_or(false, |x| x == "hip"))
        .collect();
    srcs.sort();
    let jobs: usize = std::env::var("OMRNTT_JOBS").ok().and_then(|j| j.parse().ok()).unwrap_or(8);
    let mut objs = Vec::new();
    for chunk in srcs.chunks(jobs.max(1)) {
        let mut running = Vec::new();
        for src in chunk {
            let obj = out.join(src.file_stem().unwrap()).with_extension("o");
            let child = Command::new(&hipcc)
                .args(flags)
                .arg("-c")
                .arg(src)
                .arg("-o")
                .arg(&obj)
                .spawn()
                .unwrap_or_else(|e| panic!("cannot run {hipcc}: {e}"));
            running.push((child, src.clone()));
            objs.push(obj);
        }
        for (mut child, src) in running {
            assert!(child.wait().unwrap().success(), "hipcc failed on {}", src.display());
        }
    }
    let lib = out.join("libomrntt.so");
    let status = Command::new(&hipcc)
        .args(["--offload-arch=gfx950", "-shared", "-fPIC", "-o"])
        .arg(&lib)
        .args(&objs)
        .status()
        .expect("link libomrntt.so");
    assert!(status.success(), "linking libomrntt.so failed");
    // the shipped tuning table is looked up beside the library
    let tune = manifest.join("../../tuning/mi355x.tune");
    if tune.exists() {
        std::fs::create_dir_all(out.join("tuning")).unwrap();
        std::fs::copy(&tune, out.join("tuning/mi355x.tune")).unwrap();
    }
    println!("cargo:rustc-link-search=native={}", out.display());
    println!("cargo:rustc-link-lib=dylib=omrntt");
    println!("cargo:rustc-link-arg=-Wl,-rpath,{}", out.display());
    println!("cargo:rerun-if-changed={}", csrc.display());
    println!("cargo:rerun-if-changed={}", Path::new(&manifest).join("../../../include/omrntt.h").display());
    println!("cargo:rerun-if-env-changed=HIPCC");
}
