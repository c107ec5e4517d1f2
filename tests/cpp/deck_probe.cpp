// Prints every key of an input deck as "key<TAB>kind<TAB>value" (test driver
// for a2c_amd/host/deck.hpp).
#include <cstdio>

#include "deck.hpp"

int main(int argc, char** argv) {
  try {
    a2c::Deck d = a2c::Deck::parse_file(argv[1]);
    for (const auto& kv : d.values()) {
      const a2c::DeckValue& v = kv.second;
      if (v.kind == a2c::DeckValue::INT) std::printf("%s\tint\t%lld\n", kv.first.c_str(), v.i);
      else if (v.kind == a2c::DeckValue::DBL) std::printf("%s\tdbl\t%.17g\n", kv.first.c_str(), v.d);
      else std::printf("%s\t%s\t%s\n", kv.first.c_str(), v.kind == a2c::DeckValue::STR ? "str" : "sym", v.s.c_str());
    }
  } catch (const std::exception& e) {
    std::printf("ERROR %s\n", e.what());
    return 1;
  }
  return 0;
}
