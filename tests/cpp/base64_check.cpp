// Decodes an OpenLB checkpoint with the C++ mirror's unbase64, checks the
// size header against the body, re-encodes header and body with base64 and
// prints "same" if that reproduces the file byte for byte.  Built and run by
// tests/test_host_mirror.py; needs no GPU.
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <string>

#include "olb_gfx950.h"

int main(int argc, char** argv)
{
  if (argc < 2) return 2;
  std::ifstream in(argv[1], std::ios::binary);
  const std::string text((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
  const auto head = olb::unbase64(text.substr(0, 12));
  const auto body = olb::unbase64(text.substr(12));
  std::uint64_t size = 0;
  std::memcpy(&size, head.data(), 8);
  std::printf("%llu %zu\n", (unsigned long long)size, body.size());
  const std::string again = olb::base64(head.data(), 8) + olb::base64(body.data(), body.size());
  std::printf("%s\n", again == text ? "same" : "different");
  return 0;
}
