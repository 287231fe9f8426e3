// Reads lines "chunk_cost multiple c_0 c_1 ... c_{n-1}" on stdin and prints the chunk
// starts qiw_chunk_starts returns for each, one line each (CPU test driver).
#include <iostream>
#include <sstream>
#include <string>
#include "qiw_chunks.h"

int main() {
    std::string line;
    while (std::getline(std::cin, line)) {
        std::istringstream in(line);
        int64_t chunk_cost = 0;
        int multiple = 1;
        in >> chunk_cost >> multiple;
        std::vector<int64_t> tcost;
        for (int64_t c; in >> c;) tcost.push_back(c);
        const std::vector<size_t> starts = qiw_chunk_starts(tcost, chunk_cost, multiple);
        for (size_t i = 0; i < starts.size(); ++i) std::cout << (i ? " " : "") << starts[i];
        std::cout << "\n";
    }
    return 0;
}
