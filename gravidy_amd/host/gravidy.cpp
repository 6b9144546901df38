// gravidy.cpp -- driver of the HIP backend: the same sequence as the
// reference's main() (gravidy.cpp:53-115) with Hermite4HIP as the backend.
#include "hip/Hermite4HIP.hpp"

int main(int argc, char *argv[]) {
    OptionsParser op(argc, argv);
    if (!op.check_options()) return 1;

    NbodySystem ns(op);
    if (!ns.read_input_file() || ns.n == 0) {
        std::cerr << "gravidy: no particles read from " << ns.input_filename << std::endl;
        return 1;
    }
    ns.alloc_base_attributes(0);
    ns.copy_input_data();

    Logger logger(&ns);
    NbodyUtils nu(&ns);

    Hermite4HIP h4(&ns, &logger, &nu, ns.precision, ns.dt_min, ns.dt_max);
    if (!op.checkpoint_filename.empty()) h4.set_checkpoint(op.checkpoint_filename);
    if (!op.restart_filename.empty()) h4.set_restart(op.restart_filename);
    h4.integration();

    logger.log_success(std::string("Finishing..."));
    return 0;
}
