//! Builds libparticular_amd.so for gfx950 with hipcc and links it.
//!
//! Environment:
//! * `PARTICULAR_AMD_ROOT`: root of the particular_amd sources (the directory holding
//!   `include/` and `particular_amd/csrc/`). Required unless the `prebuilt` feature is on.
//! * `PARTICULAR_AMD_LIB_DIR`: with the `prebuilt` feature, the directory that holds an already
//!   built `libparticular_amd.so` (e.g. from `python -m particular_amd.build`).
//! * `HIPCC`: the hipcc to use (default `/opt/rocm/bin/hipcc`).
//!
//! The flags are those of `particular_amd/build.py`: gfx950 only, -O3, packed FP32 instructions
//! off, and the Barnes-Hut file built without machine sinking (its walk issues the scalar loads of
//! both possible next records at the top of each step).

use std::env;
use std::path::{Path, PathBuf};
use std::process::Command;

fn hip_sources(csrc: &Path) -> Vec<PathBuf> {
    let mut v: Vec<PathBuf> = std::fs::read_dir(csrc)
        .unwrap_or_else(|e| panic!("reading {}: {e}", csrc.display()))
        .filter_map(|e| e.ok().map(|e| e.path()))
        .filter(|p| p.extension().map_or(false, |x| x == "hip"))
        .collect();
    v.sort();
    v
}

fn run(cmd: &mut Command) {
    let status = cmd.status().unwrap_or_else(|e| panic!("running {cmd:?}: {e}"));
    assert!(status.success(), "{cmd:?} failed with {status}");
}

fn main() {
    println!("cargo:rerun-if-env-changed=PARTICULAR_AMD_ROOT");
    println!("cargo:rerun-if-env-changed=PARTICULAR_AMD_LIB_DIR");
    println!("cargo:rerun-if-env-changed=HIPCC");

    if env::var_os("CARGO_FEATURE_PREBUILT").is_some() {
        let dir = env::var("PARTICULAR_AMD_LIB_DIR").expect("PARTICULAR_AMD_LIB_DIR (prebuilt)");
        println!("cargo:rustc-link-search=native={dir}");
        println!("cargo:rustc-link-lib=dylib=particular_amd");
        println!("cargo:rustc-link-arg=-Wl,-rpath,{dir}");
        return;
    }

    let root = PathBuf::from(env::var("PARTICULAR_AMD_ROOT").expect("PARTICULAR_AMD_ROOT"));
    let csrc = root.join("particular_amd").join("csrc");
    let include = root.join("include");
    let out = PathBuf::from(env::var("OUT_DIR").unwrap());
    let hipcc = env::var("HIPCC").unwrap_or_else(|_| "/opt/rocm/bin/hipcc".into());
    println!("cargo:rerun-if-changed={}", csrc.display());
    println!("cargo:rerun-if-changed={}", include.display());

    let mut objects = Vec::new();
    for src in hip_sources(&csrc) {
        let name = src.file_name().unwrap().to_string_lossy().into_owned();
        let obj = out.join(format!("{name}.o"));
        let mut cmd = Command::new(&hipcc);
        cmd.args(["--offload-arch=gfx950", "-mcode-object-version=5", "-O3", "-std=c++17"])
            .args(["-fPIC", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"])
            .arg(format!("-I{}", include.display()));
        if name == "bh.hip" {
            cmd.args(["-mllvm", "-disable-machine-sink"]);
        }
        run(cmd.arg("-c").arg(&src).arg("-o").arg(&obj));
        objects.push(obj);
    }
    let lib = out.join("libparticular_amd.so");
    run(Command::new(&hipcc)
        .args(["--offload-arch=gfx950", "-mcode-object-version=5", "-shared", "-fPIC", "-o"])
        .arg(&lib)
        .args(&objects));
    println!("cargo:rustc-link-search=native={}", out.display());
    println!("cargo:rustc-link-lib=dylib=particular_amd");
    println!("cargo:rustc-link-arg=-Wl,-rpath,{}", out.display());
}
