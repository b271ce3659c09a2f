# Generates versionutils.cpp for the reference engine build in oracle/_ref/.
# It runs CMake's own configure_file on the reference's template
# (src/engine/versionutils.cpp.in, CMakeLists.txt:155-158) with the version
# numbers the reference's CMakeLists.txt:52-59 sets; the label and licence id
# only feed the engine's log greeting.
#   cmake -DREF=<reference root> -DOUT=<output file> -P configure_version.cmake
set(FFENGINE_VERSION_MAJOR 1)
set(FFENGINE_VERSION_MINOR 8)
set(FFENGINE_VERSION_REVISION 5)
set(FFENGINE_VERSION_SUPPORT_LICENSE_TYPE "GitHub")
set(FFENGINE_VERSION_LABEL "1.8.5 GitHub Release (oracle build)")
set(FFENGINE_VERSION_SUPPORT_LICENSE_ID "oracle")
configure_file("${REF}/src/engine/versionutils.cpp.in" "${OUT}")
