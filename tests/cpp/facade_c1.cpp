// Drives the C++ facade (include/shmoai.hpp) with the reference's call sequence:
//   the C1 pipeline NTT -> rotate_vector_inplace(1) -> multiply_plain_inplace(encode(0.37))
//   -> rescale_to_next_inplace -> inverse NTT (test_phantom_ckks.cuh-style facade calls), and
//   one ct_pt_matrix_mul_wo_pre of 4 inputs by a 4 x 3 weight matrix,
// on limbs read from files, writing the resulting limbs for the Python test to compare
// with the oracle. Also checks that errors arrive as the reference's exception classes
// and that public-key encryption decrypts to its message.
// usage: facade_c1 <ct_in> <ct_out> <mm_in> <mm_out>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <vector>

#include "shmoai.hpp"

using namespace moai;

static std::vector<uint64_t> read_words(const char *path, size_t n) {
    std::vector<uint64_t> v(n);
    std::ifstream f(path, std::ios::binary);
    f.read(reinterpret_cast<char *>(v.data()), n * 8);
    if (!f) throw std::runtime_error(std::string("cannot read ") + path);
    return v;
}

static void write_words(const char *path, const std::vector<uint64_t> &v) {
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char *>(v.data()), v.size() * 8);
}

int main(int argc, char **argv) {
    if (argc != 5) {
        std::cerr << "usage: facade_c1 <ct_in> <ct_out> <mm_in> <mm_out>\n";
        return 2;
    }
    try {
        std::vector<int> bits(8, 50);
        bits.push_back(60);
        PhantomContext context(13, bits, 1);
        PhantomCKKSEncoder phantom_encoder(context);
        Encoder encoder(&context, &phantom_encoder);
        Evaluator evaluator(&context, &phantom_encoder);
        PhantomSecretKey secret_key(context, 1, 64);
        std::vector<int> steps{1};
        PhantomGaloisKey galois_keys = secret_key.create_galois_keys_from_steps(context, steps, 9);
        const size_t N = context.poly_modulus_degree(), Ql = 8;

        // C1 pipeline
        PhantomCiphertext ct(context, 2, 1);
        ct.load(read_words(argv[1], 2 * Ql * N));
        evaluator.transform_to_ntt_inplace(ct);
        evaluator.rotate_vector_inplace(ct, 1, galois_keys);
        PhantomPlaintext pt;
        encoder.encode(0.37, std::pow(2.0, 50), pt);
        evaluator.multiply_plain_inplace(ct, pt);
        evaluator.rescale_to_next_inplace(ct);
        evaluator.transform_from_ntt_inplace(ct);
        write_words(argv[2], ct.limbs());
        std::cout << "c1 pipeline: chain " << ct.chain_index() << ", " << ct.coeff_modulus_size() << " limbs\n";

        // ct-pt matmul, 4 inputs x (4 x 3) weights at scale 2^40
        std::vector<uint64_t> mm = read_words(argv[3], 4 * 2 * Ql * N);
        std::vector<PhantomCiphertext> X;
        for (int j = 0; j < 4; j++) {
            X.emplace_back(context, 2, 1);
            X.back().load(std::vector<uint64_t>(mm.begin() + j * 2 * Ql * N, mm.begin() + (j + 1) * 2 * Ql * N));
            X.back().scale() = std::pow(2.0, 40);
        }
        std::vector<std::vector<double>> W(4, std::vector<double>(3));
        for (int j = 0; j < 4; j++)
            for (int i = 0; i < 3; i++) W[j][i] = 0.01 * (j + 1) - 0.005 * i;
        std::vector<PhantomCiphertext> Y = ct_pt_matrix_mul_wo_pre(X, W, 4, 3, 4, context);
        std::vector<uint64_t> out;
        for (auto &y : Y) {
            auto l = y.limbs();
            out.insert(out.end(), l.begin(), l.end());
        }
        write_words(argv[4], out);
        std::cout << "matmul: " << Y.size() << " outputs, scale 2^" << std::log2((double)Y[0].scale()) << "\n";

        // errors arrive as the reference's exception classes
        PhantomCiphertext a(X[0]), b(X[1]);
        b.scale() = std::pow(2.0, 41);
        bool caught = false;
        try {
            evaluator.sub_inplace(a, b);  // the reference's sub checks the scales
        } catch (const std::invalid_argument &) {
            caught = true;
        }
        std::cout << (caught ? "invalid_argument: ok" : "invalid_argument: MISSING") << "\n";

        // public-key encryption through moai::Encryptor, then moai::Decryptor
        PhantomPublicKey public_key = secret_key.gen_publickey(context, 17);
        Encryptor encryptor(&context, &public_key, 100);
        Decryptor decryptor(&context, &secret_key);
        std::vector<double> v(encoder.slot_count());
        for (size_t t = 0; t < v.size(); t++) v[t] = std::sin(0.001 * (double)t);
        PhantomPlaintext pv, pd;
        encoder.encode(v, std::pow(2.0, 40), pv);
        PhantomCiphertext cv;
        encryptor.encrypt(pv, cv);
        evaluator.multiply_const_inplace(cv, 0.5);
        evaluator.rescale_to_next_inplace(cv);
        decryptor.decrypt(cv, pd);
        std::vector<double> back;
        encoder.decode(pd, back);
        double err = 0;
        for (size_t t = 0; t < v.size(); t++) err = std::max(err, std::fabs(back[t] - 0.5 * v[t]));
        std::cout << "encrypt/multiply_const/decrypt max err " << err << "\n";
        if (!caught || err > 1e-5) return 1;
        std::cout << "FACADE_OK\n";
        return 0;
    } catch (const std::exception &e) {
        std::cerr << "facade_c1: " << e.what() << "\n";
        return 1;
    }
}
